"""Continuous batching on the GPU: every request's result is the same alone and in the mix (1e-5 relative in log10 n
for predictions, 1e-5 relative for gradients), batches of different buckets run concurrently, and the captured HIP
graph of a bucket replays to the eager result every time, with other GPU work between replays."""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _backend(seed=7):
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    from ffnoml_amd.backend import GPUBackend
    config = dict(in_channels=6, out_channels=6, width=32, n_layers=2, hidden_spectral=32, hidden_vertical=32,
                  freq_mix_rank=4, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H"]))
    model = synthetic.random_model(seed, **config)
    mean = np.array([3.8, 0, 0, 0, 16, -7], np.float32)
    std = np.array([0.4, 3, 3, 3, 1.2, 2.2], np.float32)
    model.set_normalization(mean, std, np.zeros(6), np.ones(6))
    return GPUBackend(None, model=model, level_names=[f"H{i}" for i in range(6)])


def _request(rng, nz, nx, ny, z0=0.0):
    from ffnoml_amd import synthetic
    feats, zc = synthetic.atmosphere_features(1, nz, nx, ny, seed=int(rng.integers(1 << 30)))
    si = feats[0].copy()
    si[0] = 10.0 ** si[0]
    si[1:4] *= 1e3
    z = np.broadcast_to((zc + z0)[:, None, None] * 1e6, si.shape[1:]).astype(np.float32).copy()
    return si.astype(np.float32), z


def _rel(a, b):
    return float((np.abs(np.log10(a) - np.log10(b)) / np.maximum(np.abs(np.log10(b)), 1.0)).max())


def test_results_alone_and_in_the_mix():
    from ffnoml_amd.serving import ContinuousBatcher
    backend = _backend()
    rng = np.random.default_rng(0)
    shapes = [(16, 8, 8), (16, 8, 8), (16, 8, 8), (24, 16, 12), (24, 16, 12), (12, 32, 32)]
    reqs = []
    for i, shp in enumerate(shapes):
        f, z = _request(rng, *shp, z0=0.5 * (i % 2))          # same grid up to a shift: the same bucket
        op = "VJP" if i % 3 == 2 else "PREDICT"
        cot = rng.standard_normal(shp + (6,)).astype(np.float32) * 1e-12 if op == "VJP" else None
        reqs.append((op, f, z, cot))
    alone = []
    for op, f, z, cot in reqs:
        if op == "PREDICT":
            alone.append(backend.predict_local(f, z, 48e3, 48e3))
        else:
            alone.append(backend.vjp_local(f, z, 48e3, 48e3, cot))
    batcher = ContinuousBatcher({"H": backend}, slots=3, max_batch=4)
    try:
        for rep in range(2):                                   # the second pass replays captured graphs
            with cf.ThreadPoolExecutor(8) as ex:
                futs = [ex.submit(lambda a: batcher.submit(a[0], "H", a[1], a[2], 48e3, 48e3, a[3]).result(), r)
                        for r in reqs]
                mixed = [fu.result() for fu in futs]
            for (op, *_), a, m in zip(reqs, alone, mixed):
                if op == "PREDICT":
                    err = _rel(m["populations"], a)
                    print(f"pass {rep} predict rel {err:.2e}", flush=True)
                    assert err <= 1e-5
                else:
                    for k in ("features", "z_scale"):
                        err = float(np.linalg.norm(m[k] - a[k]) / np.linalg.norm(a[k]))
                        print(f"pass {rep} vjp {k} rel {err:.2e}", flush=True)
                        assert err <= 1e-5
        lat = batcher.latencies_ms()
        assert len(lat) == 2 * len(reqs)
        print(f"peak batches in flight {batcher.peak_inflight}; graphs {len(batcher.graphs)}", flush=True)
    finally:
        batcher.close()


@pytest.mark.parametrize("nz,nx,ny", [(20, 16, 16), (12, 130, 132)])
def test_graph_replays_match_eager(nz, nx, ny):
    """Replays of one captured forward, each after unrelated eager GPU work, against the eager forward; the second
    tile's horizontal DFTs run factorised along both directions (130 = 10 x 13, 132 = 11 x 12)."""
    from ffnoml_amd.serving import ContinuousBatcher
    backend = _backend(seed=9)
    rng = np.random.default_rng(5)
    f, z = _request(rng, nz, nx, ny)
    eager = backend.predict_local(f, z, 48e3, 48e3)
    batcher = ContinuousBatcher({"H": backend}, slots=1, max_batch=1)
    try:
        for i in range(4):
            got = batcher.submit("PREDICT", "H", f, z, 48e3, 48e3).result()["populations"]
            a = torch.randn(1 << 20, device="cuda")
            (a * a).sum().item()                               # eager work between replays
            err = _rel(got, eager)
            print(f"replay {i} rel {err:.2e}", flush=True)
            assert err <= 1e-6
        assert len(batcher.graphs) == 1
    finally:
        batcher.close()


def test_serving_trace_tiles_with_the_benchmark_model():
    """The serving workload's tile shapes (96 x 32^2, 128 x 64^2, 192 x 128^2; forward and VJP) with the benchmark
    model (width 64, 8 layers, 32 vertical modes) through the batcher, every result against the backend alone
    (1e-5); the smallest tile against the float64 oracle in SI units (tests/golden/wl_service.npz): log10 n 1e-3,
    the feature gradient per channel and the per-level height gradient 1e-2."""
    import hashlib
    import json
    import os
    from golden_cases import GOLDEN
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    from ffnoml_amd.backend import GPUBackend
    from ffnoml_amd.serving import ContinuousBatcher
    data = np.load(os.path.join(GOLDEN, "wl_service.npz"))
    config = json.loads(str(data["config_json"]))
    config["species"] = species(config["species"])
    _, _, mx, sx = synthetic.atmosphere(8, 128, 64, 64, seed=0)
    model = synthetic.random_model(0, **config)
    model.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    h = hashlib.sha256()
    for k, v in sorted(model.state_dict().items()):
        h.update(k.encode())
        h.update(np.ascontiguousarray(v.double().numpy()).tobytes())
    assert h.hexdigest() == str(data["state_sha256"])
    backend = GPUBackend(None, model=model, level_names=[f"L{i}" for i in range(12)])
    tiles = [(96, 32, 32), (128, 64, 64), (192, 128, 128)]
    reqs = [(op, *synthetic.si_request(*t)) for t in tiles for op in ("PREDICT", "VJP")]
    alone = []
    for op, f, z, cot in reqs:
        print(f"[service] alone {op} {z.shape}", flush=True)
        alone.append(backend.predict_local(f, z, 48e3, 48e3) if op == "PREDICT" else
                     backend.vjp_local(f, z, 48e3, 48e3, cot))
    batcher = ContinuousBatcher({"m": backend}, slots=3, max_batch=4)
    try:
        with cf.ThreadPoolExecutor(8) as ex:
            futs = [ex.submit(lambda a: batcher.submit(a[0], "m", a[1], a[2], 48e3, 48e3,
                                                       a[3] if a[0] == "VJP" else None).result(), r) for r in reqs]
            mixed = [fu.result() for fu in futs]
    finally:
        batcher.close()
    for (op, f, z, _), a, m in zip(reqs, alone, mixed):
        if op == "PREDICT":
            err = _rel(m["populations"], a)
        else:
            err = max(float(np.linalg.norm(m[k] - a[k]) / np.linalg.norm(a[k])) for k in ("features", "z_scale"))
        print(f"[service] {op} {z.shape} mix vs alone {err:.2e}", flush=True)
        assert err <= 1e-5
    stride = int(data["stride"])
    n = alone[0]                                                  # (nz, nx, ny, 12)
    log_n = np.log10(n.transpose(3, 0, 1, 2)[:, ::stride, ::stride, ::stride].astype(np.float64))
    ref = data["log10_n"].astype(np.float64)
    err = float((np.abs(log_n - ref) / np.maximum(np.abs(ref), 1.0)).max())
    vjp = alone[1]
    fb = vjp["features"][:, ::stride, ::stride, ::stride].astype(np.float64)
    fref = data["feature_bar"].astype(np.float64)
    ferr = max(float(np.linalg.norm(fb[c] - fref[c]) / np.linalg.norm(fref[c])) for c in range(6))
    zl = vjp["z_scale"].astype(np.float64).sum(axis=(1, 2))
    zerr = float(np.linalg.norm(zl - data["z_bar_level"]) / np.linalg.norm(data["z_bar_level"]))
    print(f"[service] 96x32x32 vs oracle: log10 n {err:.2e}, feature gradient {ferr:.2e}, height gradient {zerr:.2e}",
          flush=True)
    assert err <= 1e-3
    assert ferr <= 1e-2
    assert zerr <= 1e-2


def test_pinned_stage_round_trips_in_chunks():
    """The workers' pinned double buffer moves stacked arrays to the device and results back bit for bit, for arrays
    several chunks long and not a multiple of the chunk, and for a non-contiguous device tensor."""
    from ffnoml_amd.staging import PinnedStage
    rng = np.random.default_rng(3)
    stage = PinnedStage(chunk_bytes=1 << 20)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for shape in [(3, 5, 7), (6, 37, 41, 29), (12, 64, 33)]:
            arrays = [rng.standard_normal(shape).astype(np.float32) for _ in range(3)]
            t = stage.to_device(arrays, torch.device("cuda:0"), stream)
            assert torch.equal(t.cpu(), torch.from_numpy(np.stack(arrays)))
            back = stage.to_host(t.permute(0, *range(t.dim() - 1, 0, -1)), stream)
            assert np.array_equal(back, np.stack(arrays).transpose(0, *range(len(shape), 0, -1)))
    stream.synchronize()


def test_socket_service_end_to_end():
    """The protocol-v2 socket front end over the continuous batcher on the GPU: a client that speaks the wire format
    (HELLO, READY with world >= 2, PREDICT / VJP / JVP with Fortran-order float32 arrays, MEMORY, SHUTDOWN) gets the
    backend's own answers (1e-5 relative in log10 n; 1e-5 relative for the feature and height gradients and the
    tangent), two clients at once."""
    import socket
    import threading
    from ffnoml_amd.service import Service, read_array, write_array
    from ffnoml_amd.serving import ContinuousBatcher
    backend = _backend()
    batcher = ContinuousBatcher({"H": backend}, slots=2, max_batch=4)
    svc = Service(batcher, "tok", host="127.0.0.1", port=0, log=lambda s: None)
    server = threading.Thread(target=svc.serve_forever, daemon=True)
    server.start()
    rng = np.random.default_rng(5)
    nz, nx, ny = 16, 12, 10
    L = len(backend.level_names)
    errors = []

    def client(seed):
        try:
            r = np.random.default_rng(seed)
            f, z = _request(r, nz, nx, ny)
            cot = r.standard_normal((nz, nx, ny, L)).astype(np.float32) * 1e-12
            tan = r.standard_normal((6, nz, nx, ny)).astype(np.float32) * f.std(axis=(1, 2, 3))[:, None, None, None]
            s = socket.create_connection(("127.0.0.1", svc.port), timeout=60)
            rd = s.makefile("rb", buffering=0)
            s.sendall(b"HELLO tok\n")
            ready = rd.readline().decode().split()
            assert ready[0] == "READY" and int(ready[2]) >= 2
            s.sendall(f"PREDICT H {nz} {nx} {ny} {L} 48000.0 48000.0\n".encode())
            write_array(s, f)
            write_array(s, z)
            assert rd.readline().decode().split() == ["OK", "PREDICT", str(nz), str(nx), str(ny), str(L)]
            pops = read_array(s, (nz, nx, ny, L))
            assert _rel(pops, backend.predict_local(f, z, 48e3, 48e3)) <= 1e-5
            s.sendall(f"VJP H {nz} {nx} {ny} {L} 48000.0 48000.0\n".encode())
            for a in (f, z, cot):
                write_array(s, a)
            assert rd.readline().decode().split()[:2] == ["OK", "VJP"]
            fb, zb = read_array(s, (6, nz, nx, ny)), read_array(s, (nz, nx, ny))
            ref = backend.vjp_local(f, z, 48e3, 48e3, cot)
            for got, want in ((fb, ref["features"]), (zb, ref["z_scale"])):
                assert float(np.linalg.norm(got - want) / np.linalg.norm(want)) <= 1e-5
            s.sendall(f"JVP H {nz} {nx} {ny} {L} 48000.0 48000.0\n".encode())
            for a in (f, z, tan):
                write_array(s, a)
            assert rd.readline().decode().split()[:2] == ["OK", "JVP"]
            jt = read_array(s, (nz, nx, ny, L))
            want = backend.jvp_local(f, z, 48e3, 48e3, tan)
            assert float(np.linalg.norm(jt - want) / np.linalg.norm(want)) <= 1e-5
            s.sendall(b"MEMORY\n")
            assert rd.readline().decode().split() == ["OK", "MEMORY", "1"]
            rd.readline()
            s.close()
        except Exception as e:          # reported by the main thread
            errors.append(e)

    try:
        threads = [threading.Thread(target=client, args=(seed,)) for seed in (11, 12)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=180)
        assert not errors, errors
        s = socket.create_connection(("127.0.0.1", svc.port), timeout=60)
        rd = s.makefile("rb", buffering=0)
        s.sendall(b"HELLO tok\n")
        rd.readline()
        s.sendall(b"SHUTDOWN\n")
        assert rd.readline() == b"BYE\n"
        s.close()
        server.join(timeout=10)
        assert not server.is_alive()
    finally:
        svc.stopping.set()
        batcher.close()
