"""The C-ABI boundary without a GPU: symbols, registry, hooks, JSON and the
circuit text format, mirroring tests/unit/test_scheduler.c and
tests/unit/test_qgtl_backend.c of the reference.  Nothing here executes a
circuit on a device."""
import ctypes
import json
import os
import re
import threading

import numpy as np
import pytest

import moonlab_amd as m
from conftest import GOLDEN, ROOT

INCLUDE = os.path.join(ROOT, "include")


ROUTE_HEADER = "moonlab_gpu_route.h"   # the reference's GPU seam: libmoonlab_mi355x_route.so


def declared_symbols():
    names = set()
    for f in os.listdir(INCLUDE):
        if not f.endswith(".h") or f == ROUTE_HEADER:
            continue
        text = open(os.path.join(INCLUDE, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for mm in re.finditer(r"MOONLAB_API\s+[^;{]*?\b(moonlab_\w+)\s*\(", text):
            names.add(mm.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = m.lib()
    names = declared_symbols()
    assert len(names) > 80
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and nothing exported under the moonlab_ prefix is left undeclared
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", m.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split()[-1].startswith("moonlab_")}
    assert exported == names, sorted(exported ^ names)


def test_route_library_exports_the_seam():
    """libmoonlab_mi355x_route.so exports exactly the seam that
    include/moonlab_gpu_route.h declares (the symbols of the reference's
    state_gpu.cu / gpu_route_cpu_unavailable.c), and the engine library none
    of them."""
    import subprocess
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(INCLUDE, ROUTE_HEADER)).read(), flags=re.S)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    names = {mm.group(1) for mm in re.finditer(r"MOONLAB_API\s+[^;{]*?\b(\w+)\s*\(", text)}
    assert {"quantum_state_create_gpu", "qsim_gpu_route_cnot", "moonlab_cuda_state_free"} <= names
    route = os.path.join(os.path.dirname(m.LIB_PATH), "libmoonlab_mi355x_route.so")
    out = subprocess.run(["nm", "-D", "--defined-only", route], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split()[1] in ("T", "W")}
    assert exported == names, sorted(exported ^ names)
    main = subprocess.run(["nm", "-D", "--defined-only", m.LIB_PATH], capture_output=True, text=True).stdout
    assert not names & {ln.split()[-1] for ln in main.splitlines()}


def test_routed_reference_without_device():
    """The reference built with the seam in place of its CPU stub
    (oracle/Makefile ref_routed): the seam's strong definitions replace
    state.c's weak ones, nothing is left undefined, and with no GPU the
    contract of the reference's tests/unit/test_gpu_sync_contract.c holds:
    quantum_state_create_gpu fails loudly (NOT_SUPPORTED, NULL out), the sync
    calls are no-ops on a CPU state and reject NULL."""
    import subprocess
    routed = os.path.join(ROOT, "oracle", "_ref", "libmoonlab_ref_routed.so")
    if not os.path.exists(routed):
        pytest.skip("oracle/_ref/libmoonlab_ref_routed.so not built (make -C oracle ref_routed)")
    out = subprocess.run(["nm", "-D", routed], capture_output=True, text=True).stdout
    kinds = {ln.split()[-1]: ln.split()[-2] for ln in out.splitlines() if len(ln.split()) >= 2}
    for sym in ("quantum_state_create_gpu", "quantum_state_sync_to_host", "quantum_state_sync_from_host",
                "qsim_gpu_route_apply_1q_matrix", "qsim_gpu_route_apply_2q_matrix", "moonlab_cuda_state_free"):
        assert kinds.get(sym) == "T", (sym, kinds.get(sym))
    undefined = subprocess.run(["nm", "-D", "--undefined-only", routed], capture_output=True, text=True).stdout
    plugin = os.path.join(os.path.dirname(m.LIB_PATH), "libmoonlab_mi355x_plugin.so")
    pdef = subprocess.run(["nm", "-D", "--defined-only", plugin], capture_output=True, text=True).stdout
    provided = {ln.split()[-1] for ln in pdef.splitlines()}
    left = [ln.split()[-1] for ln in undefined.splitlines()
            if ("qsim_" in ln or "quantum_" in ln or "moonlab_" in ln) and " w " not in ln]
    assert left and set(left) <= provided, sorted(set(left) - provided)   # the engine, from the plugin
    if m.device_count() > 0:
        pytest.skip("a GPU is visible: the device path is tested by tests/test_gpu_engine.py")
    L = ctypes.CDLL(routed)
    L.quantum_state_create_gpu.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
    p = ctypes.c_void_p(1)
    assert L.quantum_state_create_gpu(4, ctypes.byref(p)) == -7 and not p.value
    L.quantum_state_sync_to_host.argtypes = [ctypes.c_void_p]
    assert L.quantum_state_sync_to_host(None) == -6
    L.ref_state_new.restype = ctypes.c_void_p
    L.ref_state_new.argtypes = [ctypes.c_int]
    L.ref_state_delete.argtypes = [ctypes.c_void_p]
    s = L.ref_state_new(3)
    assert L.quantum_state_sync_to_host(s) == 0
    L.ref_state_delete(s)


def test_binding_declares_every_symbol_it_calls():
    """ctypes passes an undeclared argument as a C int, which truncates a
    64-bit handle: every engine symbol the Python binding calls must carry
    argtypes (the declaration table in moonlab_amd/__init__.py)."""
    src = open(os.path.join(ROOT, "moonlab_amd", "__init__.py")).read()
    called = set(re.findall(r"(?:lib\(\)|_lib|\bL)\.(moonlab_\w+)\(", src))
    declared = set(re.findall(r'"(moonlab_\w+)": \(', src))
    assert called, "no calls found"
    assert called <= declared, sorted(called - declared)
    L = m.lib()
    for name in called:
        assert getattr(L, name).argtypes is not None, name


def test_hea_pack_layout():
    """HeaPack: descriptors point into one flat angle buffer at each
    instance's offset, in the layout moonlab_mi355x_hea_t declares."""
    from moonlab_amd import workloads
    inst = workloads.hea_batch(40, 7)
    p = m.HeaPack(inst)
    assert p.count == 40 and p.angles.dtype == np.float64
    for i, (n, L, ang, h) in enumerate(inst):
        assert p._arr["num_qubits"][i] == n and p._arr["num_layers"][i] == L and p._arr["h"][i] == h
        assert p._arr["angles"][i] == p.angles.ctypes.data + 8 * int(p.offsets[i])
        assert np.array_equal(p.angles[p.offsets[i]:p.offsets[i] + 2 * n * L], np.asarray(ang).reshape(-1))
    assert p.offsets[-1] + 2 * inst[-1][0] * inst[-1][1] == p.angles.size


def test_job_introspection_and_errors():
    L = m.lib()
    assert not L.moonlab_job_create(0)
    assert not L.moonlab_job_create(33)
    j = m.Job(4)
    assert (j.num_qubits, j.num_gates, j.num_shots, j.num_workers) == (4, 0, 0, 1)
    j.add_gate(m.GateType.H, 0)
    assert j.num_gates == 1
    j.set_num_shots(100).set_num_workers(4)
    assert (j.num_shots, j.num_workers) == (100, 4)
    assert L.moonlab_job_set_num_shots(j.handle, -1) == m.MOONLAB_SCHED_BAD_ARG
    assert L.moonlab_job_set_num_workers(j.handle, 0) == m.MOONLAB_SCHED_BAD_ARG
    assert L.moonlab_job_add_gate(j.handle, int(m.GateType.RX), 0, -1, None) == m.MOONLAB_SCHED_BAD_ARG
    assert L.moonlab_job_add_gate(None, 4, 0, -1, None) == m.MOONLAB_SCHED_BAD_ARG
    # no validation of qubit indices at add time (scheduler.c:99-121)
    assert L.moonlab_job_add_gate(j.handle, 4, 99, -1, None) == 0
    j0 = m.Job(2)
    rc, _ = j0.run_raw()
    assert rc == m.MOONLAB_SCHED_BAD_ARG  # zero shots
    assert L.moonlab_scheduler_run_mpi(j0.handle, None, None) == m.MOONLAB_SCHED_NOT_BUILT


def test_job_json_byte_identical_to_reference():
    shots = json.load(open(os.path.join(GOLDEN, "shots_ref.json")))
    rnd10 = np.load(os.path.join(GOLDEN, "rnd10_gates.npy"))
    circuits = {"bell": [(4, 0, -1, 0.0), (10, 1, 0, 0.0)],
                "ghz": [(4, 0, -1, 0.0), (10, 1, 0, 0.0), (10, 2, 1, 0.0)],
                "rnd10": [(int(g["type"]), int(g["target"]), int(g["control"]), float(g["theta"]))
                          for g in rnd10]}
    for name, gl in circuits.items():
        s = shots[name]
        j = m.Job(s["n"]).set_num_shots(s["shots"]).set_num_workers(s["workers"]).set_rng_seed(s["seed"])
        for t, tg, c, th in gl:
            j.add_gate(t, tg, c, th if t in (7, 8, 9) else None)
        assert j.to_json() == s["json"], name
    # size probe and truncation contract
    L = m.lib()
    need = L.moonlab_job_to_json(j.handle, None, 0)
    small = ctypes.create_string_buffer(16)
    assert L.moonlab_job_to_json(j.handle, small, 16) == need
    assert small.value.decode() == s["json"][:15]


def test_registry_default_and_stub_dispatch():
    names = m.list_backends()
    assert names[0] == "simulator"
    L = m.lib()
    sim = L.moonlab_find_backend(b"simulator")
    assert sim and sim.contents.name == b"simulator"
    hits = []

    @m.BACKEND_EXECUTE_FN
    def stub(job, out, ctx):
        hits.append(1)
        n = L.moonlab_job_num_shots(job)
        libc = ctypes.CDLL(None)
        libc.malloc.restype = ctypes.c_void_p
        libc.calloc.restype = ctypes.c_void_p
        out.contents.num_qubits = L.moonlab_job_num_qubits(job)
        out.contents.total_shots = n
        out.contents.num_workers_used = 1
        out.contents.outcomes = ctypes.cast(libc.malloc(8 * n), ctypes.POINTER(ctypes.c_uint64))
        out.contents.worker_seconds = ctypes.cast(libc.calloc(1, 8), ctypes.POINTER(ctypes.c_double))
        for i in range(n):
            out.contents.outcomes[i] = 0xA5
        return 0

    be = m._Backend(b"test-stub", stub, None, b"stub")
    assert L.moonlab_register_backend(ctypes.byref(be)) == 0
    assert L.moonlab_find_backend(b"test-stub")
    n0 = L.moonlab_num_backends()
    assert L.moonlab_register_backend(ctypes.byref(be)) == 0  # same name replaces
    assert L.moonlab_num_backends() == n0
    j = m.Job(2).add_gate(4, 0).set_num_shots(64).set_backend("test-stub")
    assert j.backend == "test-stub"
    r = j.execute()
    assert len(hits) == 1 and r.total_shots == 64 and set(r.outcomes) == {0xA5}
    assert L.moonlab_unregister_backend(b"test-stub") == 0
    assert not L.moonlab_find_backend(b"test-stub")
    assert L.moonlab_unregister_backend(b"test-stub") == m.MOONLAB_SCHED_BACKEND_NOT_FOUND
    j.set_backend("does-not-exist")
    rc, _ = j.run_raw()
    assert rc == m.MOONLAB_SCHED_BACKEND_NOT_FOUND
    j.set_backend(None)
    assert j.backend is None
    assert L.moonlab_register_backend(None) == m.MOONLAB_SCHED_BAD_ARG


def test_registry_capacity_is_sixteen():
    L = m.lib()

    @m.BACKEND_EXECUTE_FN
    def nop(job, out, ctx):
        return m.MOONLAB_SCHED_INTERNAL

    base = L.moonlab_num_backends()
    keep = []
    added = []
    try:
        for i in range(32):
            name = f"cap-{i}".encode()
            be = m._Backend(name, nop, None, None)
            keep.append((be, name))
            rc = L.moonlab_register_backend(ctypes.byref(be))
            if rc != 0:
                assert rc == m.MOONLAB_SCHED_OOM
                break
            added.append(name)
        assert base + len(added) == 16
    finally:
        for name in added:
            L.moonlab_unregister_backend(name)


def test_completion_hook_and_request_context():
    L = m.lib()
    seen = []

    @m.COMPLETION_HOOK_FN
    def hook(job, res, backend, ctx):
        seen.append((backend.decode(), res.contents.total_shots,
                     L.moonlab_scheduler_current_tenant_id()))

    @m.BACKEND_EXECUTE_FN
    def ok(job, out, ctx):
        libc = ctypes.CDLL(None)
        libc.calloc.restype = ctypes.c_void_p
        n = L.moonlab_job_num_shots(job)
        out.contents.total_shots = n
        out.contents.num_workers_used = 1
        out.contents.outcomes = ctypes.cast(libc.calloc(n, 8), ctypes.POINTER(ctypes.c_uint64))
        out.contents.worker_seconds = ctypes.cast(libc.calloc(1, 8), ctypes.POINTER(ctypes.c_double))
        return 0

    @m.BACKEND_EXECUTE_FN
    def bad(job, out, ctx):
        return m.MOONLAB_SCHED_INTERNAL

    b1 = m._Backend(b"hook-ok", ok, None, None)
    b2 = m._Backend(b"hook-bad", bad, None, None)
    L.moonlab_register_backend(ctypes.byref(b1))
    L.moonlab_register_backend(ctypes.byref(b2))
    try:
        assert L.moonlab_scheduler_set_completion_hook(hook, None) == 0
        L.moonlab_scheduler_set_request_context(b"acme-corp", b"req-42")
        m.Job(2).set_num_shots(8).set_backend("hook-ok").execute()
        assert seen == [("hook-ok", 8, b"acme-corp")]
        rc, _ = m.Job(2).set_num_shots(8).set_backend("hook-bad").run_raw()
        assert rc == m.MOONLAB_SCHED_INTERNAL and len(seen) == 1  # not fired on failure
        # request context is per thread
        other = []
        t = threading.Thread(target=lambda: other.append(L.moonlab_scheduler_current_tenant_id()))
        t.start()
        t.join()
        assert other == [None]
        L.moonlab_scheduler_set_request_context(None, None)
        L.moonlab_scheduler_set_completion_hook(m.COMPLETION_HOOK_FN(), None)
        m.Job(2).set_num_shots(8).set_backend("hook-ok").execute()
        assert len(seen) == 1
    finally:
        L.moonlab_scheduler_set_completion_hook(m.COMPLETION_HOOK_FN(), None)
        L.moonlab_unregister_backend(b"hook-ok")
        L.moonlab_unregister_backend(b"hook-bad")


def test_invalid_qubit_surfaces_as_internal_before_any_device_work():
    """A bad index is an INTERNAL error at run time (scheduler.c:159-166);
    the engine rejects it before touching the GPU."""
    j = m.Job(2).add_gate(m.GateType.H, 5).set_num_shots(4)
    rc, _ = j.run_raw()
    assert rc == m.MOONLAB_SCHED_INTERNAL
    j = m.Job(2).add_gate(m.GateType.CNOT, 1, 1).set_num_shots(4).set_backend(None)
    rc, _ = j.run_raw()
    assert rc == m.MOONLAB_SCHED_INTERNAL


def test_circuit_api_validation_and_text_format():
    L = m.lib()
    assert not L.moonlab_qgtl_circuit_create(0)
    assert not L.moonlab_qgtl_circuit_create(33)
    c = m.Circuit(2)
    assert c.add_gate_raw(m.GateType.H, 2) == m.MOONLAB_QGTL_BAD_ARG
    assert c.add_gate_raw(m.GateType.CNOT, 0, 0) == m.MOONLAB_QGTL_BAD_ARG
    assert c.add_gate_raw(m.GateType.RX, 0) == m.MOONLAB_QGTL_BAD_ARG
    L.moonlab_qgtl_circuit_free(None)
    shots = json.load(open(os.path.join(GOLDEN, "shots_ref.json")))
    rnd10 = np.load(os.path.join(GOLDEN, "rnd10_gates.npy"))
    c = m.Circuit(10)
    for g in rnd10:
        t = int(g["type"])
        c.add_gate(t, int(g["target"]), int(g["control"]), float(g["theta"]) if t in (7, 8, 9) else None)
    text = c.serialize()
    assert text == shots["qgtl_rnd10"]["text"]
    c2 = m.Circuit.deserialize(text)
    assert c2.serialize() == text and c2.num_gates == c.num_gates
    # parser edge cases (moonlab_qgtl_backend.c:414-506)
    ok = "# c\n\n   NUM_QUBITS 3\n# gate\nH 0\n  CNOT 1 0\nRZ 2 0.25\n"
    assert m.Circuit.deserialize(ok).num_gates == 3
    for bad in ["", "H 0\n", "NUM_QUBITS 0\n", "NUM_QUBITS 33\n", "NUM_QUBITS 2\nFOO 0\n",
                "NUM_QUBITS 2\nH 2\n", "NUM_QUBITS 2\nRX 0\n", "NUM_QUBITS 2\nCNOT 0 0\n",
                "NUM_QUBITS 2\n" + "H" * 300 + "\n"]:
        with pytest.raises(m.QGTLError):
            m.Circuit.deserialize(bad)
    st = ctypes.c_int(0)
    payload = b"NUM_QUBITS 2\nH 0\x00\n"
    assert not L.moonlab_qgtl_circuit_deserialize(payload, len(payload), ctypes.byref(st))
    assert st.value == m.MOONLAB_QGTL_BAD_ARG
    # size query and undersized buffer
    need = ctypes.c_size_t(0)
    assert L.moonlab_qgtl_circuit_serialize(c._h, None, 0, ctypes.byref(need)) == 0
    assert need.value == len(text)
    buf = ctypes.create_string_buffer(10)
    w = ctypes.c_size_t(0)
    assert L.moonlab_qgtl_circuit_serialize(c._h, buf, 10, ctypes.byref(w)) == m.MOONLAB_QGTL_OOM


def test_circuit_save_load_roundtrip(tmp_path):
    c = m.Circuit(3).add_gate(m.GateType.H, 0).add_gate(m.GateType.RY, 1, theta=0.1234567890123)
    p = str(tmp_path / "c.txt").encode()
    assert m.lib().moonlab_qgtl_circuit_save(c._h, p) == 0
    st = ctypes.c_int(-1)
    h = m.lib().moonlab_qgtl_circuit_load(p, ctypes.byref(st))
    assert h and st.value == 0
    assert m.Circuit(_handle=h).serialize() == c.serialize()


def test_no_device_fails_loudly():
    """Without a visible GPU the engine refuses to run: there is no CPU path."""
    if m.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(m.EngineError, match="no HIP device"):
        m.StateVector(4)
    rc, _ = m.Job(2).add_gate(4, 0).set_num_shots(4).run_raw()
    assert rc == m.MOONLAB_SCHED_INTERNAL


def test_control_frame_grammar_without_device():
    """Frames the control plane answers before any execution
    (control_plane.c:418-425, 590-684): HEALTH, verb and limit errors,
    deserialize and qubit-cap rejections, with the reference's codes."""
    hf = m.handle_frame
    assert hf(b"HEALTH\n") == (m.CONTROL_OK, b"OK alive\n")
    text = m.Circuit(2).add_gate(m.GateType.H, 0).serialize().encode()
    cases = [
        (b"HELLO\n", m.CONTROL_PROTOCOL, b"ERR -404 unknown verb\n"),
        (b"CIRCUIT x\n", m.CONTROL_PROTOCOL, b"ERR -404 expected CIRCUIT <N>\n"),
        (b"SHOTS 5\n", m.CONTROL_PROTOCOL, b"ERR -404 expected SHOTS <shots> <bytes>\n"),
        (b"SHOTS 0 10\n", m.CONTROL_BAD_ARG, b"ERR -401 shots out of range\n"),
        (b"SHOTS %d 10\n" % ((1 << 20) + 1), m.CONTROL_BAD_ARG, b"ERR -401 shots out of range\n"),
        (b"CIRCUIT 0\n", m.CONTROL_BAD_ARG, b"ERR -401 body bytes out of range\n"),
        (b"CIRCUIT %d\n" % ((1 << 22) + 1), m.CONTROL_BAD_ARG, b"ERR -401 body bytes out of range\n"),
        (b"CIRCUIT 100\n" + text, m.CONTROL_PROTOCOL, b"ERR -404 short body\n"),
        (b"CIRCUIT 7\nFOO 0\n\n", m.CONTROL_REJECTED, b"ERR -405 deserialize\n"),
        (b"HEALTH", m.CONTROL_PROTOCOL, b"ERR -404 no verb line\n"),
    ]
    for frame, rc, resp in cases:
        assert hf(frame) == (rc, resp), frame
    big = b"NUM_QUBITS 25\nH 0\n"
    assert hf(b"CIRCUIT %d\n" % len(big) + big) == (m.CONTROL_BAD_ARG, b"ERR -401 qubit count exceeds server cap\n")
    # a valid frame with no server to run it is refused at execution
    assert hf(b"CIRCUIT %d\n" % len(text) + text) == (m.CONTROL_REJECTED, b"ERR -405 execute\n")


def test_control_frame_fuzz_without_device():
    """Mutated and random frames never crash the parser and always get one
    well-formed reply line (the reference fuzzes its control plane the same
    way; a frame it would execute is refused here for want of a server)."""
    rng = np.random.default_rng(20261202)
    text = m.Circuit(3).add_gate(m.GateType.H, 0).add_gate(m.GateType.CNOT, 1, 0).serialize().encode()
    ext = b"# moonlab-circuit v1\nNUM_QUBITS 2\nP 0 nan\nU3 1 0.5 inf 0.25\n"
    seeds = [b"HEALTH\n", b"CIRCUIT %d\n" % len(text) + text, b"SHOTS 16 %d\n" % len(text) + text]
    # non-finite extension angles are refused at deserialize, before execution
    assert m.handle_frame(b"CIRCUIT %d\n" % len(ext) + ext) == (m.CONTROL_REJECTED, b"ERR -405 deserialize\n")
    codes = {m.CONTROL_OK, m.CONTROL_BAD_ARG, m.CONTROL_PROTOCOL, m.CONTROL_REJECTED}
    for i in range(3000):
        base = bytearray(seeds[i % 3])
        k = int(rng.integers(0, 3))
        if k == 0:      # flip bytes
            for _ in range(int(rng.integers(1, 6))):
                base[int(rng.integers(0, len(base)))] = int(rng.integers(0, 256))
        elif k == 1:    # truncate
            base = base[: int(rng.integers(0, len(base) + 1))]
        else:           # random bytes
            base = bytearray(rng.integers(0, 256, int(rng.integers(0, 200)), dtype=np.uint8).tobytes())
        rc, resp = m.handle_frame(bytes(base))
        assert rc in codes, (bytes(base), rc)
        if rc == m.CONTROL_OK:
            assert resp == b"OK alive\n"
        else:
            assert resp.startswith(b"ERR %d " % rc) and resp.endswith(b"\n") and resp.count(b"\n") == 1


def test_control_frame_parsers_under_sanitizers(tmp_path):
    """tests/frame_fuzz.cpp: 200k fuzzed frames through the frame and circuit
    parsers built with AddressSanitizer + UBSan (host code only)."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "frame_fuzz")
    csrc = os.path.join(ROOT, "moonlab_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I" + os.path.join(ROOT, "include"), "-I" + csrc, os.path.join(ROOT, "tests", "frame_fuzz.cpp"),
                    os.path.join(csrc, "control_frame.cpp"), os.path.join(csrc, "qgtl_circuit.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("fuzzed 200000 frames"), r.stdout


def test_circuit_extension_gates_text_and_validation():
    """moonlab_mi355x_circuit_add_gate_ex: codes 14..24 recorded with their
    qubits and parameters, written and read back as extension lines; v1
    gates keep their lines; bad arity, repeats and codes are refused; a code
    above 13 added through moonlab_qgtl_add_gate still executes as the
    reference's UNSUPPORTED (checked before any device work)."""
    c = m.Circuit(4).add_gate(m.GateType.H, 0)
    c.add_gate_ex(14, [1]).add_gate_ex(16, [2], [0.25]).add_gate_ex(17, [3], [0.1, 0.2, 0.3])
    c.add_gate_ex(18, [0, 1], [1.5]).add_gate_ex(21, [2, 3], [-0.75]).add_gate_ex(22, [1, 0])
    c.add_gate_ex(23, [3, 0, 1]).add_gate_ex(24, [2, 0, 3])
    text = c.serialize()
    lines = text.splitlines()
    assert lines[2] == "H 0" and lines[3] == "SDG 1" and lines[4] == "P 2 0.25"
    assert lines[5] == "U3 3 0.10000000000000001 0.20000000000000001 0.29999999999999999"
    assert lines[6] == "CPHASE 0 1 1.5" and lines[9] == "CCX 3 0 1" and lines[10] == "CSWAP 2 0 3"
    c2 = m.Circuit.deserialize(text)
    assert c2.serialize() == text and c2.num_gates == 9
    for bad in [(13, [0], []), (25, [0], []), (14, [0, 1], []), (16, [0], []), (17, [0], [1.0]),
                (23, [0, 0, 1], []), (18, [0, 4], [1.0]), (16, [0], [float("nan")]),
                (17, [0], [0.1, float("inf"), 0.3]), (21, [0, 1], [float("-inf")])]:
        with pytest.raises(m.QGTLError):
            m.Circuit(4).add_gate_ex(*bad)
    for bad in ["NUM_QUBITS 2\nCCX 0 1\n", "NUM_QUBITS 2\nP 0\n", "NUM_QUBITS 3\nCSWAP 0 1 1\n",
                "NUM_QUBITS 2\nP 0 nan\n", "NUM_QUBITS 2\nU3 1 0.5 inf 0.1\n", "NUM_QUBITS 2\nCRZ 0 1 -inf\n"]:
        with pytest.raises(m.QGTLError):
            m.Circuit.deserialize(bad)
    raw = m.Circuit(2)
    assert raw.add_gate_raw(14, 0) == 0          # recorded, as the reference records it
    with pytest.raises(m.QGTLError, match="-303"):
        raw.execute(return_probabilities=True)
