"""Hosts the MI355X plugin library inside the reference's own scheduler.

Run in a fresh process by the tests (the reference library is loaded with
RTLD_GLOBAL so the plugin's job-API references resolve to it, which must not
leak into the test runner's own process):

    python tests/plugin_host.py register   # registry only, no device
    python tests/plugin_host.py run        # one job on "simulator" and on "mi355x"

Prints one JSON object.  The reference library is oracle/_ref's build of the
reference C (src/distributed/scheduler.c and its closure).
"""
import ctypes
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmoonlab_ref.so")
PLUGIN_SO = os.path.join(ROOT, "moonlab_amd", "lib", "libmoonlab_mi355x_plugin.so")


class _Backend(ctypes.Structure):   # moonlab_backend_t (scheduler.h:240-245)
    _fields_ = [("name", ctypes.c_char_p), ("execute", ctypes.c_void_p),
                ("ctx", ctypes.c_void_p), ("description", ctypes.c_char_p)]


class _Results(ctypes.Structure):   # moonlab_job_results_t (scheduler.h:51-57)
    _fields_ = [("num_qubits", ctypes.c_int), ("total_shots", ctypes.c_int),
                ("outcomes", ctypes.POINTER(ctypes.c_uint64)), ("num_workers_used", ctypes.c_int),
                ("worker_seconds", ctypes.POINTER(ctypes.c_double))]


def load():
    ref = ctypes.CDLL(REF_SO, mode=ctypes.RTLD_GLOBAL)
    plug = ctypes.CDLL(PLUGIN_SO)
    vp = ctypes.c_void_p
    ref.moonlab_job_create.restype = vp
    ref.moonlab_job_create.argtypes = [ctypes.c_int]
    ref.moonlab_job_free.argtypes = [vp]
    ref.moonlab_job_add_gate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    for f in ("moonlab_job_set_num_shots", "moonlab_job_set_num_workers"):
        getattr(ref, f).argtypes = [vp, ctypes.c_int]
    ref.moonlab_job_set_rng_seed.argtypes = [vp, ctypes.c_uint64]
    ref.moonlab_job_set_backend.argtypes = [vp, ctypes.c_char_p]
    ref.moonlab_scheduler_run.argtypes = [vp, ctypes.POINTER(_Results)]
    ref.moonlab_job_results_free.argtypes = [ctypes.POINTER(_Results)]
    ref.moonlab_find_backend.restype = ctypes.POINTER(_Backend)
    ref.moonlab_find_backend.argtypes = [ctypes.c_char_p]
    ref.moonlab_list_backends.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    return ref, plug


def register(ref, plug):
    rc = plug.moonlab_mi355x_register_backend()
    names = (ctypes.c_char_p * 16)()
    k = ref.moonlab_list_backends(names, 16)
    be = ref.moonlab_find_backend(b"mi355x")
    execute = ctypes.cast(plug.moonlab_mi355x_backend_execute, ctypes.c_void_p).value
    return {"rc": rc, "backends": [names[i].decode() for i in range(k)],
            "found": bool(be), "execute_is_plugin": bool(be) and be.contents.execute == execute}


def run(ref, plug):
    from moonlab_amd import workloads
    out = register(ref, plug)
    n, shots, workers, seed = 12, 3000, 3, 0x5EED1234
    g = workloads.random_layer_circuit(n, 10, 4242, h_layer=True)
    job = ref.moonlab_job_create(n)
    for row in g:
        t = int(row["type"])
        th = ctypes.c_double(float(row["theta"]))
        ref.moonlab_job_add_gate(job, t, int(row["target"]), int(row["control"]),
                                 ctypes.byref(th) if t in (7, 8, 9) else None)
    ref.moonlab_job_set_num_shots(job, shots)
    ref.moonlab_job_set_num_workers(job, workers)
    ref.moonlab_job_set_rng_seed(job, seed)
    got = {}
    for name in (b"simulator", b"mi355x"):
        ref.moonlab_job_set_backend(job, name)
        r = _Results()
        rc = ref.moonlab_scheduler_run(job, ctypes.byref(r))
        got[name.decode()] = {"rc": rc, "outcomes": [int(r.outcomes[i]) for i in range(r.total_shots)]
                              if rc == 0 else []}
        if rc == 0:
            ref.moonlab_job_results_free(ctypes.byref(r))
    ref.moonlab_job_free(job)
    out.update(got)
    return out


if __name__ == "__main__":
    ref, plug = load()
    what = sys.argv[1] if len(sys.argv) > 1 else "register"
    print(json.dumps(register(ref, plug) if what == "register" else run(ref, plug)))
