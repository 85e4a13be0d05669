"""Out-of-process use of the device-resident integrator through AMUSE's
socket message protocol: the Python side of amuse_amd/lib/g6_worker.

The channel follows src/amuse/rfi/channel.py: the proxy listens on a local
port, starts the worker as `g6_worker PORT HOST false` and accepts its
connection (SocketChannel.start, :2211-2291); messages are SocketMessage's
(:1720-1954); function ids are |crc32(name)| read as signed int32
(src/amuse/rfi/core.py:368-417); the parameters of one dtype travel as blocks
of call_count values in parameter order, the int32 result first among the
outputs (src/amuse/rfi/tools/create_c.py:1081-1210, core.py:665-731).
Functions and parameters are ph4's (src/amuse_ph4/interface.py,
src/amuse/community/interface/gd.py).
"""
from __future__ import annotations

import os
import socket
import subprocess
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WORKER = os.path.join(HERE, "lib", "g6_worker")

HEADER_INTS = 11
ERROR_FLAG = 256


def function_id(name: str) -> int:
    """AMUSE's id of a legacy function: crc32 of the name as a signed
    32-bit integer, absolute value (core.py:376, :389-417)."""
    x = zlib.crc32(name.encode("ascii"))
    return abs(x - ((x & 0x80000000) << 1))


# name: (int inputs, double inputs, int outputs, double outputs), each a list
# of parameter names in order; every function returns an int32
SPECS = {
    "initialize_code": ([], [], [], []),
    "cleanup_code": ([], [], [], []),
    "commit_parameters": ([], [], [], []),
    "recommit_parameters": ([], [], [], []),
    "get_eps2": ([], [], [], ["epsilon_squared"]),
    "set_eps2": ([], ["epsilon_squared"], [], []),
    "get_eta": ([], [], [], ["eta"]),
    "set_eta": ([], ["eta"], [], []),
    "get_begin_time": ([], [], [], ["time"]),
    "set_begin_time": ([], ["time"], [], []),
    "get_initial_timestep_fac": ([], [], [], ["fac"]),
    "set_initial_timestep_fac": ([], ["fac"], [], []),
    "get_initial_timestep_limit": ([], [], [], ["limit"]),
    "set_initial_timestep_limit": ([], ["limit"], [], []),
    "get_initial_timestep_median": ([], [], [], ["median"]),
    "set_initial_timestep_median": ([], ["median"], [], []),
    "get_gpu": ([], [], ["use_gpu"], []),
    "set_gpu": (["use_gpu"], [], [], []),
    "get_gpu_id": ([], [], ["gpu_id"], []),
    "set_gpu_id": (["gpu_id"], [], [], []),
    "get_force_precision": ([], [], ["precision"], []),
    "set_force_precision": (["precision"], [], [], []),
    "new_particle": (["id"], ["mass", "x", "y", "z", "vx", "vy", "vz", "radius"],
                     ["index_of_the_particle"], []),
    "delete_particle": (["index_of_the_particle"], [], [], []),
    "set_state": (["index_of_the_particle"], ["mass", "x", "y", "z", "vx", "vy", "vz", "radius"],
                  [], []),
    "set_mass": (["index_of_the_particle"], ["mass"], [], []),
    "set_radius": (["index_of_the_particle"], ["radius"], [], []),
    "set_position": (["index_of_the_particle"], ["x", "y", "z"], [], []),
    "set_velocity": (["index_of_the_particle"], ["vx", "vy", "vz"], [], []),
    "commit_particles": ([], [], [], []),
    "recommit_particles": ([], [], [], []),
    "synchronize_model": ([], [], [], []),
    "evolve_model": ([], ["time"], [], []),
    "get_time": ([], [], [], ["time"]),
    "get_block_steps": ([], [], ["block_steps"], []),
    "get_total_steps": ([], [], ["total_steps"], []),
    "get_number_of_particles": ([], [], ["number_of_particles"], []),
    "get_total_mass": ([], [], [], ["mass"]),
    "get_kinetic_energy": ([], [], [], ["kinetic_energy"]),
    "get_potential_energy": ([], [], [], ["potential_energy"]),
    "get_state": (["index_of_the_particle"], [],
                  [], ["mass", "x", "y", "z", "vx", "vy", "vz", "radius"]),
    "get_mass": (["index_of_the_particle"], [], [], ["mass"]),
    "get_radius": (["index_of_the_particle"], [], [], ["radius"]),
    "get_position": (["index_of_the_particle"], [], [], ["x", "y", "z"]),
    "get_velocity": (["index_of_the_particle"], [], [], ["vx", "vy", "vz"]),
    "get_particle_timestep": (["index_of_the_particle"], [], [], ["timestep"]),
}


class Message:
    """One SocketMessage: header fields and the typed arrays."""

    def __init__(self, call_id=0, function_id=0, call_count=1, ints=(), longs=(), floats=(),
                 doubles=(), booleans=(), strings=(), units=(), error=False):
        self.call_id, self.function_id, self.call_count = call_id, function_id, call_count
        self.ints = np.asarray(ints, dtype=np.int32).ravel()
        self.longs = np.asarray(longs, dtype=np.int64).ravel()
        self.floats = np.asarray(floats, dtype=np.float32).ravel()
        self.doubles = np.asarray(doubles, dtype=np.float64).ravel()
        self.booleans = np.asarray(booleans, dtype=np.int8).ravel()
        self.strings = list(strings)
        self.units = np.asarray(units, dtype=np.float64).ravel()
        self.error = bool(error)

    def encode(self) -> bytes:
        flags = np.array([0, int(self.error), int(self.units.size > 0), 0], dtype=np.int8)
        header = np.array([self.call_id, self.function_id, self.call_count, self.ints.size,
                           self.longs.size, self.floats.size, self.doubles.size,
                           self.booleans.size, len(self.strings), self.units.size],
                          dtype=np.int32)
        parts = [flags.tobytes(), header.tobytes(), self.ints.tobytes(), self.longs.tobytes(),
                 self.floats.tobytes(), self.doubles.tobytes(), self.booleans.tobytes()]
        if self.strings:
            parts.append(np.array([len(s.encode("utf-8")) for s in self.strings],
                                  dtype=np.int32).tobytes())
            parts.append(b"".join(s.encode("utf-8") + b"\0" for s in self.strings))
        parts.append(self.units.tobytes())
        return b"".join(parts)

    @classmethod
    def read(cls, recv) -> "Message":
        """Reads one message with recv(nbytes) -> bytes."""
        head = recv(4 * HEADER_INTS)
        flags = np.frombuffer(head, dtype=np.int8, count=4)
        if flags[0] != 0:
            raise IOError("endianness of the message does not match")
        h = np.frombuffer(head, dtype=np.int32)
        ints = np.frombuffer(recv(4 * h[4]), dtype=np.int32)
        longs = np.frombuffer(recv(8 * h[5]), dtype=np.int64)
        floats = np.frombuffer(recv(4 * h[6]), dtype=np.float32)
        doubles = np.frombuffer(recv(8 * h[7]), dtype=np.float64)
        booleans = np.frombuffer(recv(h[8]), dtype=np.int8)
        strings = []
        if h[9] > 0:
            lengths = np.frombuffer(recv(4 * h[9]), dtype=np.int32)
            chars = recv(int(lengths.sum()) + int(h[9]))
            at = 0
            for n in lengths:
                strings.append(chars[at:at + n].decode("utf-8"))
                at += int(n) + 1
        units = np.frombuffer(recv(8 * h[10]), dtype=np.float64)
        return cls(int(h[1]), int(h[2]), int(h[3]), ints, longs, floats, doubles, booleans,
                   strings, units, error=bool(flags[1]))


class WorkerError(RuntimeError):
    pass


class Ph4Worker:
    """ph4's interface in a separate process (amuse_amd/lib/g6_worker)."""

    def __init__(self, path: str = WORKER, host: str = "127.0.0.1"):
        if not os.path.exists(path):
            raise WorkerError(f"{path} is not built; run amuse_amd.build")
        server = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        server.bind((host, 0))
        server.listen(1)
        server.settimeout(60)
        port = server.getsockname()[1]
        self.process = subprocess.Popen([path, str(port), host, "false"])
        try:
            self.sock, _ = server.accept()
        finally:
            server.close()
        self.sock.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
        self.sock.settimeout(None)
        self.call_id = 0

    def _recv(self, n: int) -> bytes:
        chunks = []
        while n > 0:
            b = self.sock.recv(min(n, 1 << 20))
            if not b:
                raise WorkerError("lost connection to the worker")
            chunks.append(b)
            n -= len(b)
        return b"".join(chunks)

    def send(self, msg: Message) -> Message:
        self.sock.sendall(msg.encode())
        return Message.read(self._recv)

    def call(self, name: str, *args, call_count: int | None = None):
        """Calls `name` with its inputs in parameter order (scalars or arrays
        of one length); returns (result, outputs) with arrays for array
        calls and scalars otherwise."""
        int_in, dbl_in, int_out, dbl_out = SPECS[name]
        names = int_in + dbl_in
        if len(args) != len(names):
            raise TypeError(f"{name} takes {names}")
        # the inputs in their declared order: ints, then doubles
        order = list(SPECS_ORDER[name])
        vals = dict(zip(order, args))
        arrays = [np.atleast_1d(np.asarray(vals[p])) for p in order]
        n = call_count or max([a.size for a in arrays] + [1])
        ints = [np.broadcast_to(np.asarray(vals[p], dtype=np.int32), (n,)) for p in int_in]
        dbls = [np.broadcast_to(np.asarray(vals[p], dtype=np.float64), (n,)) for p in dbl_in]
        self.call_id += 1
        msg = Message(self.call_id, function_id(name), n,
                      ints=np.concatenate(ints) if ints else (),
                      doubles=np.concatenate(dbls) if dbls else ())
        rep = self.send(msg)
        if rep.error:
            raise WorkerError(f"{name}: {' '.join(rep.strings)}")
        if rep.call_id != self.call_id or rep.function_id != msg.function_id:
            raise WorkerError(f"{name}: reply does not match the call")
        ri = rep.ints.reshape(1 + len(int_out), n)
        rd = rep.doubles.reshape(len(dbl_out), n) if dbl_out else np.zeros((0, n))
        scalar = all(np.ndim(vals[p]) == 0 for p in order) and call_count is None
        outs = {}
        for k, p in enumerate(int_out):
            outs[p] = int(ri[1 + k, 0]) if scalar else ri[1 + k].copy()
        for k, p in enumerate(dbl_out):
            outs[p] = float(rd[k, 0]) if scalar else rd[k].copy()
        result = int(ri[0, 0]) if scalar else ri[0].copy()
        return result, outs

    def stop(self) -> int:
        """The _stop_worker call (id 0, core.py:955-959), then the channel
        closes and the process ends."""
        if self.sock is None:
            return self.process.wait()
        self.call_id += 1
        self.send(Message(self.call_id, 0, 1))
        self.sock.close()
        self.sock = None
        return self.process.wait(timeout=60)

    def __del__(self):
        try:
            if getattr(self, "sock", None) is not None:
                self.stop()
        except Exception:
            pass


# parameter order of each function's inputs as declared (gd.py / ph4's
# interface.py); SPECS lists them split by dtype
SPECS_ORDER = {name: (spec[0] + spec[1]) for name, spec in SPECS.items()}
SPECS_ORDER["new_particle"] = ["mass", "x", "y", "z", "vx", "vy", "vz", "radius", "id"]
