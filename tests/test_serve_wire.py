"""CPU: the served form's wire format (JSON jobs and results), the job struct
the native server takes, its result digest (pinned by the published FNV-1a
64 test vectors), and the server's refusals that need no device."""
import ctypes as C
import json

import pytest

from neurocorrelation_amd import serve as SV
from neurocorrelation_amd._lib import FNV_OFFSET, JobResult, NCError, lib


def fnv1a64(data, h=FNV_OFFSET):
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_digest_is_fnv1a_64():
    # published FNV-1a 64-bit test vectors
    assert lib().nc_digest(b"", 0, FNV_OFFSET) == 0xCBF29CE484222325
    assert lib().nc_digest(b"a", 1, FNV_OFFSET) == 0xAF63DC4C8601EC8C
    assert lib().nc_digest(b"foobar", 6, FNV_OFFSET) == 0x85944171F73967E8
    blob = bytes(range(256)) * 3
    assert lib().nc_digest(blob, len(blob), FNV_OFFSET) == fnv1a64(blob)
    # continued over pieces = over the whole
    h = lib().nc_digest(blob[:100], 100, FNV_OFFSET)
    assert lib().nc_digest(blob[100:], len(blob) - 100, h) == fnv1a64(blob)


def test_digest_of_arrays_hashes_potentials_weights_firings_in_order():
    import numpy as np
    V = np.array([-70.0, 1.5], np.float32)
    w = np.array([0.25, -0.5, 1.0], np.float32)
    f = np.array([3, 0], np.uint32)
    want = fnv1a64(V.tobytes() + w.tobytes() + f.tobytes())
    assert SV.digest_arrays(V, w, f) == want
    assert SV.summarize(V, w, f) == {"spikes": 3, "digest": "%016x" % want}


def test_job_json_round_trip():
    specs = SV.served_specs(3, seed=4)
    specs[1]["params"] = {"dt": 0.5, "learning_rate": 0.5, "pre_factor": 0.13, "post_factor": 0.3,
                          "pre_trace_decay": 0.75, "post_trace_decay": 0.65}
    specs[2]["drive"]["bg_prob"] = 0.001
    for i, sp in enumerate(specs):
        line = SV.job_to_json(100 + i, sp)
        job_id, back = SV.job_from_json(line)
        assert job_id == 100 + i
        assert back["steps"] == sp["steps"] and back["wiring"] == sp["wiring"]
        assert {k: back["drive"][k] for k in SV.DRIVE_KEYS} == {k: sp["drive"][k] for k in SV.DRIVE_KEYS}
        assert back.get("params") == sp.get("params")
        assert back["drive"].get("bg_prob") == sp["drive"].get("bg_prob")
        assert SV.job_to_json(job_id, back) == line


@pytest.mark.parametrize("bad", [
    '{"id": 1, "steps": 10, "wiring": {}}',
    '{"id": 1, "steps": 0, "wiring": %s, "drive": %s}',
    '{"steps": 10, "wiring": %s, "drive": %s}',
])
def test_malformed_jobs_are_refused(bad):
    sp = SV.served_specs(1, seed=5)[0]
    if "%s" in bad:
        bad = bad % (json.dumps(sp["wiring"]), json.dumps(sp["drive"]))
    with pytest.raises(ValueError):
        SV.job_from_json(bad)


def test_job_struct_carries_the_spec():
    sp = SV.served_specs(1, seed=6)[0]
    j = SV.job_struct(42, sp)
    assert j.id == 42 and j.steps == sp["steps"]
    assert j.wiring.n_neurons == sp["wiring"]["n_neurons"] and j.wiring.seed == sp["wiring"]["seed"]
    assert abs(j.wiring.radius - sp["wiring"]["radius"]) < 1e-6 and j.wiring.spatial_order == 1
    assert j.drive.n_inputs == sp["drive"]["n_inputs"] and j.drive.seed == sp["drive"]["seed"]
    assert abs(j.params.dt - 1.0) < 1e-9 and abs(j.params.post_factor - 0.30) < 1e-7


def test_result_record_layout():
    # the C struct: 8 + 4 + 4 + 4 + 4 + 8 + 8 + 8 + 4 * 8 bytes
    assert C.sizeof(JobResult) == 80


def test_server_without_a_device_is_refused_cleanly():
    try:
        srv = SV.Server(max_neurons=1000, max_synapses=10_000)
    except NCError as e:
        assert e.code in (3, 7)  # NC_ERR_HIP, NC_ERR_NO_DEVICE
    else:
        srv.close()
        pytest.skip("a GPU is present")
    # null arguments are codes, not crashes
    assert lib().nc_server_create(None, None) != 0
    assert lib().nc_server_submit(None, None) != 0
    assert lib().nc_server_destroy(None) == 0
