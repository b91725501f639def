"""The plugin library registers into the reference's own backend registry
(src/distributed/scheduler.c:263-330), in a fresh process, without a device."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmoonlab_ref.so")


def host(what):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "plugin_host.py"), what],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_plugin_registers_in_reference_registry():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference checkout)")
    out = host("register")
    assert out["rc"] == 0
    assert "simulator" in out["backends"] and "mi355x" in out["backends"]
    assert out["found"] and out["execute_is_plugin"]


def test_plugin_leaves_job_api_to_the_host():
    """The plugin defines no job / registry entry point of its own."""
    plugin = os.path.join(ROOT, "moonlab_amd", "lib", "libmoonlab_mi355x_plugin.so")
    und = subprocess.run(["nm", "-D", "--undefined-only", plugin], capture_output=True, text=True).stdout
    dfn = subprocess.run(["nm", "-D", "--defined-only", plugin], capture_output=True, text=True).stdout
    want = {"moonlab_job_num_qubits", "moonlab_job_num_gates", "moonlab_job_num_shots",
            "moonlab_job_num_workers", "moonlab_job_to_json", "moonlab_register_backend"}
    assert want <= {l.split()[-1] for l in und.splitlines() if l.strip()}
    defined = {l.split()[-1] for l in dfn.splitlines() if l.strip()}
    assert not {d for d in defined if d.startswith(("moonlab_job_", "moonlab_scheduler_", "moonlab_qgtl_"))}
    assert "moonlab_mi355x_backend_execute" in defined and "moonlab_mi355x_register_backend" in defined
