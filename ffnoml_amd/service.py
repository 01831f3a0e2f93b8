"""Socket service speaking the reference's wire protocol v2 (ffno_fsdp_service.py:22-45, :156-343), backed by the
single-GPU continuous batcher instead of an FSDP world.

Protocol (unchanged for the client): one ASCII command line, raw little-endian float32 arrays in Fortran order.
  HELLO <token>                                   -> READY 2 <slots> <nmodels> FULL_SHARD_H_SLAB
  PREDICT|VJP|JVP <name> nz nx ny levels dx dy    -> features (6,nz,nx,ny), z (nz,nx,ny) [, cotangent or tangent]
                                                  <- OK <OP> nz nx ny levels, then the result arrays
  MEMORY                                          <- OK MEMORY 1, then "host 0 peak reserved total"
  SHUTDOWN                                        <- BYE (the service stops)
The READY line's world field carries the batcher's execution slots (at least 2): the Julia client checks
world >= 2 for its full-shard service and the tag it accepts for predict/VJP (FSDPService.jl:175-182). Several clients
may connect at once; their requests meet in the batcher. A failed request is answered "ERROR <message>" and the
connection carries on (the reference's batched service answers unsupported operations the same way,
ffno_batched_service.py:72-77). A JVP answers the population tangent (nz, nx, ny, levels), as the reference's
service does (ffno_fsdp_service.py:281-297).
"""
import argparse
import importlib
import os
import socket
import sys
import threading
from pathlib import Path

import numpy as np

PROTOCOL_VERSION = 2


def read_exact(stream, count):
    buf = bytearray(count)
    view = memoryview(buf)
    off = 0
    while off < count:
        got = stream.recv_into(view[off:])
        if got == 0:
            raise EOFError("client disconnected during array transfer")
        off += got
    return bytes(buf)


def read_array(stream, shape):
    count = int(np.prod(shape))
    return np.frombuffer(read_exact(stream, count * 4), dtype="<f4").reshape(shape, order="F").copy()


def write_array(stream, value):
    stream.sendall(np.asfortranarray(value, dtype="<f4").tobytes(order="F"))


def memory_line(device):
    import torch
    if device is None or not torch.cuda.is_available():
        return f"{socket.gethostname()} 0 0 0 0"
    _, total = torch.cuda.mem_get_info(device)
    peak = torch.cuda.max_memory_allocated(device)
    torch.cuda.reset_peak_memory_stats(device)
    return f"{socket.gethostname()} 0 {peak} {torch.cuda.memory_reserved(device)} {total}"


class Service:
    """Accepts clients on a listening socket; each client gets a thread that parses its commands and waits on the
    batcher's futures. `batcher` needs submit(op, model, features, z, dx, dy, cotangent) -> Future, `slots`, and
    `backends` (name -> object with level_names)."""

    def __init__(self, batcher, token, host="0.0.0.0", port=0, device=None, log=print):
        self.batcher = batcher
        self.token = token
        self.device = device
        self.log = log
        self.listener = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.listener.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.listener.bind((host, int(port)))
        self.listener.listen(16)
        self.port = self.listener.getsockname()[1]
        self.stopping = threading.Event()
        self.threads = []

    def serve_forever(self):
        self.log(f"FFNO_SERVICE_LISTENING port={self.port} slots={self.batcher.slots} "
                 f"models={','.join(sorted(self.batcher.backends))}")
        self.listener.settimeout(0.2)
        while not self.stopping.is_set():
            try:
                conn, address = self.listener.accept()
            except socket.timeout:
                continue
            t = threading.Thread(target=self._client, args=(conn, address), daemon=True)
            t.start()
            self.threads.append(t)
        self.listener.close()

    def _client(self, conn, address):
        conn.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
        reader = conn.makefile("rb", buffering=0)
        try:
            hello = reader.readline().decode("utf-8").strip().split()
            if len(hello) != 2 or hello[0] != "HELLO" or hello[1] != self.token:
                conn.sendall(b"ERROR invalid handshake\n")
                return
            slots = max(2, int(self.batcher.slots))
            conn.sendall(f"READY {PROTOCOL_VERSION} {slots} {len(self.batcher.backends)} FULL_SHARD_H_SLAB\n".encode())
            self.log(f"FFNO_SERVICE_CLIENT_CONNECTED address={address}")
            while True:
                line = reader.readline()
                if not line:
                    return
                command = line.decode("utf-8").strip().split()
                if not command:
                    continue
                op = command[0]
                if op == "SHUTDOWN":
                    conn.sendall(b"BYE\n")
                    self.stopping.set()
                    return
                if op == "MEMORY":
                    conn.sendall(f"OK MEMORY 1\n{memory_line(self.device)}\n".encode())
                    continue
                if op not in ("PREDICT", "VJP", "JVP") or len(command) != 8:
                    conn.sendall(f"ERROR invalid command: {' '.join(command)}\n".encode())
                    return                                  # the stream cannot be trusted after a bad command
                _, name, nz, nx, ny, levels, dx, dy = command
                nz, nx, ny, levels = map(int, (nz, nx, ny, levels))
                # the request's arrays are on the wire whatever the answer: read them to stay in step
                features = read_array(conn, (6, nz, nx, ny))
                z = read_array(conn, (nz, nx, ny))
                extra = None
                if op == "VJP":
                    extra = read_array(conn, (nz, nx, ny, levels))
                elif op == "JVP":
                    extra = read_array(conn, (6, nz, nx, ny))
                try:
                    backend = self.batcher.backends[name]
                    if levels != len(backend.level_names):
                        raise ValueError("request level count differs from the model")
                    result = self.batcher.submit(op, name, features, z, float(dx), float(dy), extra).result()
                except Exception as e:                      # answered, not fatal: other clients carry on
                    conn.sendall(f"ERROR {type(e).__name__}: {e}\n".replace("\n", " ").strip().encode() + b"\n")
                    continue
                conn.sendall(f"OK {op} {nz} {nx} {ny} {levels}\n".encode())
                if op == "PREDICT":
                    write_array(conn, result["populations"])
                elif op == "JVP":
                    write_array(conn, result["population_tangent"])
                else:
                    write_array(conn, result["features"])
                    write_array(conn, result["z_scale"])
                self.log(f"FFNO_SERVICE_COMMAND_OK operation={op} model={name} shape={nz}x{nx}x{ny}")
        except (EOFError, ConnectionError):
            return
        finally:
            reader.close()
            conn.close()


def load_manifest(path):
    try:
        import tomllib
    except ImportError:                                      # Python 3.10
        import tomli as tomllib
    with open(path, "rb") as f:
        return tomllib.load(f)


def main(argv=None):
    """python -m ffnoml_amd.service manifest.toml: the reference's manifest ([service] port, token, bind_host;
    [[models]] name, checkpoint_path, factory_module, factory_name, level_names, factory_kwargs)."""
    ap = argparse.ArgumentParser()
    ap.add_argument("manifest")
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--max-batch", type=int, default=8)
    args = ap.parse_args(argv)
    cfg = load_manifest(Path(args.manifest).resolve())
    from .backend import GPUBackend
    from .serving import ContinuousBatcher
    backends = {}
    for m in sorted(cfg["models"], key=lambda item: item["name"]):
        backends[str(m["name"])] = GPUBackend(
            m["checkpoint_path"], factory_module=m.get("factory_module", "ffnoml_amd.factory"),
            factory_name=m.get("factory_name", "create_ffno3d"), level_names=m["level_names"],
            factory_kwargs=m.get("factory_kwargs", {}), name=str(m["name"]))
    batcher = ContinuousBatcher(backends, slots=args.slots, max_batch=args.max_batch)
    svc = cfg["service"]
    service = Service(batcher, svc["token"], host=svc.get("bind_host", "0.0.0.0"), port=int(svc["port"]),
                      device=next(iter(backends.values())).device, log=lambda s: print(s, flush=True))
    try:
        service.serve_forever()
    finally:
        batcher.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
