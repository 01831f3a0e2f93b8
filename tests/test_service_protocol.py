"""The socket service's wire protocol (the reference's v2: ffno_fsdp_service.py:22-45, :156-343) on the CPU, with a
stand-in batcher, as the reference's Julia suite tests its client against a stand-in server: handshake, Fortran-order
arrays both ways, MEMORY, errors that answer without killing the service, JVP, SHUTDOWN."""
import socket
import threading
from concurrent.futures import Future

import numpy as np
import pytest

from ffnoml_amd.service import Service, read_array, write_array


class StandInBackend:
    level_names = ("a", "b", "c")


class StandInBatcher:
    """Returns deterministic arrays computed from the request, so the test can check what went over the wire."""
    slots = 2
    backends = {"H": StandInBackend()}

    def __init__(self):
        self.calls = []

    def submit(self, op, model, features, z, dx, dy, cotangent=None):
        self.calls.append((op, model, features.shape, z.shape, dx, dy))
        f = Future()
        if features[0].min() <= 0:
            f.set_exception(ValueError("FFNO request temperature must be positive"))
            return f
        nz, nx, ny = z.shape
        if op == "PREDICT":
            pops = np.stack([features[0] * (i + 1) + z for i in range(3)], -1)
            f.set_result({"populations": pops})
        elif op == "JVP":
            f.set_result({"population_tangent": np.stack([cotangent[i] + z for i in range(3)], -1)})
        else:
            f.set_result({"features": features * 2.0, "z_scale": z + cotangent.sum(-1)})
        return f


@pytest.fixture()
def service():
    svc = Service(StandInBatcher(), "tok", host="127.0.0.1", port=0, log=lambda s: None)
    t = threading.Thread(target=svc.serve_forever, daemon=True)
    t.start()
    yield svc
    svc.stopping.set()
    t.join(timeout=5)


def _connect(svc, token="tok"):
    s = socket.create_connection(("127.0.0.1", svc.port), timeout=10)
    r = s.makefile("rb", buffering=0)
    s.sendall(f"HELLO {token}\n".encode())
    return s, r


def test_handshake_predict_vjp_memory_shutdown(service):
    s, r = _connect(service)
    ready = r.readline().decode().split()
    assert ready[:2] == ["READY", "2"] and int(ready[2]) >= 2 and ready[3] == "1" and ready[4] == "FULL_SHARD_H_SLAB"
    rng = np.random.default_rng(0)
    nz, nx, ny = 5, 4, 3
    feats = rng.random((6, nz, nx, ny)).astype(np.float32) + 0.5
    z = rng.random((nz, nx, ny)).astype(np.float32)
    s.sendall(f"PREDICT H {nz} {nx} {ny} 3 48000.0 48000.0\n".encode())
    write_array(s, feats)
    write_array(s, z)
    assert r.readline().decode().split() == ["OK", "PREDICT", str(nz), str(nx), str(ny), "3"]
    pops = read_array(s, (nz, nx, ny, 3))
    assert np.allclose(pops[..., 1], feats[0] * 2 + z)
    cot = rng.random((nz, nx, ny, 3)).astype(np.float32)
    s.sendall(f"VJP H {nz} {nx} {ny} 3 48000.0 48000.0\n".encode())
    write_array(s, feats)
    write_array(s, z)
    write_array(s, cot)
    assert r.readline().decode().split()[:2] == ["OK", "VJP"]
    assert np.allclose(read_array(s, (6, nz, nx, ny)), feats * 2)
    assert np.allclose(read_array(s, (nz, nx, ny)), z + cot.sum(-1), atol=1e-6)
    s.sendall(b"MEMORY\n")
    assert r.readline().decode().split() == ["OK", "MEMORY", "1"]
    assert len(r.readline().decode().split()) == 5
    s.sendall(b"SHUTDOWN\n")
    assert r.readline() == b"BYE\n"
    s.close()


def test_errors_answer_and_the_connection_carries_on(service):
    s, r = _connect(service)
    r.readline()
    nz, nx, ny = 2, 2, 2
    bad = np.zeros((6, nz, nx, ny), np.float32)          # T = 0: the backend's ValueError
    s.sendall(f"PREDICT H {nz} {nx} {ny} 3 1.0 1.0\n".encode())
    write_array(s, bad)
    write_array(s, np.zeros((nz, nx, ny), np.float32))
    line = r.readline().decode()
    assert line.startswith("ERROR ValueError") and "positive" in line
    # JVP: features, z, then the feature tangent (6, nz, nx, ny); answered with the population tangent
    # (nz, nx, ny, levels), as the reference's service does (ffno_fsdp_service.py:281-297)
    tangent = np.arange(6 * nz * nx * ny, dtype=np.float32).reshape(6, nz, nx, ny)
    zz = np.full((nz, nx, ny), 0.5, np.float32)
    s.sendall(f"JVP H {nz} {nx} {ny} 3 1.0 1.0\n".encode())
    write_array(s, bad + 1)
    write_array(s, zz)
    write_array(s, tangent)
    assert r.readline().decode().split() == ["OK", "JVP", str(nz), str(nx), str(ny), "3"]
    got = read_array(s, (nz, nx, ny, 3))
    assert np.array_equal(got, np.stack([tangent[i] + zz for i in range(3)], -1))
    good = bad + 1
    s.sendall(f"PREDICT H {nz} {nx} {ny} 3 1.0 1.0\n".encode())
    write_array(s, good)
    write_array(s, np.zeros((nz, nx, ny), np.float32))
    assert r.readline().decode().startswith("OK PREDICT")
    read_array(s, (nz, nx, ny, 3))
    # a wrong level count is answered
    s.sendall(f"PREDICT H {nz} {nx} {ny} 4 1.0 1.0\n".encode())
    write_array(s, good)
    write_array(s, np.zeros((nz, nx, ny), np.float32))
    assert r.readline().decode().startswith("ERROR")
    s.close()


def test_bad_token_is_refused(service):
    s, r = _connect(service, token="wrong")
    assert r.readline().decode().startswith("ERROR")
    s.close()


def test_client_disconnect_mid_request_leaves_the_service_up(service):
    s, r = _connect(service)
    r.readline()
    s.sendall(b"PREDICT H 4 4 4 3 1.0 1.0\n")
    s.sendall(np.zeros(10, np.float32).tobytes())      # truncated array, then gone
    s.close()
    s2, r2 = _connect(service)
    assert r2.readline().decode().startswith("READY")
    s2.close()
