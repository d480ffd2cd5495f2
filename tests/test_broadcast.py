"""The interactive driver (world_amd/bin/broadcast) speaks the reference's
WebSocket protocol (src/broadcast.cpp:24-113, parsed by client/world.js:
587-668).  The client here is written against RFC 6455 directly (raw
sockets), and the opening handshake's Sec-WebSocket-Accept is checked
against Python's hashlib SHA-1, plus the RFC's own worked example (section
1.3), so the server's SHA-1 and base64 are pinned independently.

The protocol tests need no GPU (the server builds a simulation only on
"initialize"); the stream test compares the initialize / run messages with
the Python engine advanced over the same steps."""
import base64
import hashlib
import os
import re
import socket
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

BIN = Path(__file__).resolve().parents[1] / "world_amd" / "bin" / "broadcast"
GUID = "258EAFA5-E914-47DA-95CA-C5AB0DC85B11"

CONFIG = """
simulation:
  dt: 0.0002
  collision_grid_size_x: 100
  collision_grid_size_y: 100
  collision_grid_size_z: 70
  particles_per_cell: 8
  density_grid_size_x: 16
  density_grid_size_y: 16
  density_grid_size_z: 8
domain:
  x_min: -1.0
  x_max: 1.0
  y_min: -1.0
  y_max: 1.0
  z_min: -1.0
  z_max: 0.4
physics:
  gravity: 9.81
  particle_radius: 0.01
  max_force: 2000.0
  floor_z: -1.0
  ceiling_z: 0.4
  wall_x_min: -1.0
  wall_x_max: 1.0
  wall_y_min: -1.0
  wall_y_max: 1.0
  restitution: 0.3
  friction: 0.5
materials:
  masses: [0.0, 0.08]
initialization:
  type: cube
  x0: -0.05
  y0: -0.05
  z0: -0.96
  cube_length_x: 0.1
  cube_length_y: 0.1
  cube_length_z: 0.1
  jitter: 0.25
  seed: 12345
  particle_material: 1
rendering:
  num_triangles: 12
drivers:
  port: 0
  steps_per_frame: 5
"""

needs_bin = pytest.mark.skipif(not BIN.exists(), reason="world_amd/bin/broadcast is not built")


class Server:
    def __init__(self, tmp_path):
        cfg = tmp_path / "world.yaml"
        cfg.write_text(CONFIG)
        self.proc = subprocess.Popen([str(BIN), str(cfg)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                     text=True, bufsize=1)
        line = self.proc.stdout.readline()
        m = re.match(r"Listening on port (\d+)", line)
        assert m, line
        self.port = int(m.group(1))

    def expect_line(self, text):
        for _ in range(20):
            line = self.proc.stdout.readline()
            if text in line:
                return line
        raise AssertionError(f"server never printed {text!r}")

    def close(self):
        self.proc.kill()
        self.proc.wait(timeout=10)


@pytest.fixture
def server(tmp_path):
    s = Server(tmp_path)
    yield s
    s.close()


class Client:
    def __init__(self, port, key=None):
        self.sock = socket.create_connection(("127.0.0.1", port), timeout=60)
        self.key = key or base64.b64encode(os.urandom(16)).decode()
        self.sock.sendall((f"GET /world HTTP/1.1\r\nHost: 127.0.0.1:{port}\r\nUpgrade: websocket\r\n"
                           f"Connection: Upgrade\r\nSec-WebSocket-Key: {self.key}\r\n"
                           f"Sec-WebSocket-Version: 13\r\n\r\n").encode())
        resp = b""
        while b"\r\n\r\n" not in resp:
            chunk = self.sock.recv(1024)
            assert chunk, "connection closed during the handshake"
            resp += chunk
        head = resp.split(b"\r\n\r\n")[0].decode()
        self.status = head.split("\r\n")[0]
        self.headers = {k.strip().lower(): v.strip() for k, v in
                        (l.split(":", 1) for l in head.split("\r\n")[1:])}

    def frame(self, opcode, payload, fin=True, mask=True):
        b0 = (0x80 if fin else 0) | opcode
        n = len(payload)
        if n < 126:
            hdr = bytes([b0, (0x80 if mask else 0) | n])
        elif n < 65536:
            hdr = bytes([b0, (0x80 if mask else 0) | 126]) + struct.pack(">H", n)
        else:
            hdr = bytes([b0, (0x80 if mask else 0) | 127]) + struct.pack(">Q", n)
        if mask:
            key = os.urandom(4)
            payload = bytes(c ^ key[i & 3] for i, c in enumerate(payload))
            hdr += key
        self.sock.sendall(hdr + payload)

    def text(self, s):
        self.frame(1, s.encode())

    def _recv(self, n):
        buf = b""
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise ConnectionError("closed")
            buf += chunk
        return buf

    def recv(self):
        b0, b1 = self._recv(2)
        assert not (b1 & 0x80), "server frames are unmasked"
        n = b1 & 0x7F
        if n == 126:
            n = struct.unpack(">H", self._recv(2))[0]
        elif n == 127:
            n = struct.unpack(">Q", self._recv(8))[0]
        return b0 & 0x0F, bool(b0 & 0x80), self._recv(n)

    def close(self):
        self.sock.close()


def accept_of(key):
    return base64.b64encode(hashlib.sha1((key + GUID).encode()).digest()).decode()


@needs_bin
def test_handshake_matches_rfc_and_hashlib(server):
    c = Client(server.port, key="dGhlIHNhbXBsZSBub25jZQ==")  # RFC 6455 section 1.3
    assert c.status == "HTTP/1.1 101 Switching Protocols"
    assert c.headers["upgrade"].lower() == "websocket"
    assert c.headers["sec-websocket-accept"] == "s3pPLMBiTxaQ9kYGzzhZRbK+xOo="
    c.close()
    for _ in range(4):  # keys of other lengths' padding go through the same SHA-1 / base64
        c = Client(server.port)
        assert c.headers["sec-websocket-accept"] == accept_of(c.key)
        c.close()


@needs_bin
def test_protocol_control_frames_and_errors(server):
    c = Client(server.port)
    c.frame(9, b"are you there")  # ping -> pong with the same payload
    assert c.recv() == (10, True, b"are you there")
    c.text("hello")
    assert "Unknown message: hello" in server.expect_line("Unknown message")
    c.frame(1, b"unk", fin=False)  # a fragmented text message
    c.frame(0, b"nown", fin=True)
    assert "Received message: unknown" in server.expect_line("Received message: unknown")
    c.text("run")  # before any "initialize": no simulation to advance
    server.expect_line("no simulation yet")
    c.frame(8, struct.pack(">H", 1000))  # close -> close echoed
    op, fin, payload = c.recv()
    assert op == 8 and payload == struct.pack(">H", 1000)
    c.close()
    # The next client is served; an unmasked client frame is a protocol error.
    c = Client(server.port)
    c.frame(1, b"initialize", mask=False)
    op, fin, payload = c.recv()
    assert op == 8 and payload == struct.pack(">H", 1002)
    c.close()
    c = Client.__new__(Client)
    c.sock = socket.create_connection(("127.0.0.1", server.port), timeout=60)
    c.sock.sendall(b"GET / HTTP/1.1\r\nHost: x\r\n\r\n")  # no Sec-WebSocket-Key
    assert c.sock.recv(1024).startswith(b"HTTP/1.1 400")
    c.close()


@needs_bin
@pytest.mark.gpu
def test_initialize_and_run_streams(server, tmp_path):
    from world_amd._capi import load_config
    from world_amd.sim import ParticleDynamics

    cfg = load_config(str(tmp_path / "world.yaml"))
    sim = ParticleDynamics(cfg)
    n = sim.n
    ref0 = sim.state()[:, :3].astype(np.float32).copy()

    c = Client(server.port)
    c.text("initialize")
    msgs = [c.recv() for _ in range(6)]
    assert all(op == 2 and fin for op, fin, _ in msgs)
    assert [len(p) for _, _, p in msgs] == [4, 12, 4, 4, 24, 12 * n]
    assert struct.unpack("<i", msgs[0][2])[0] == n == 125
    assert struct.unpack("<3i", msgs[1][2]) == (100, 100, 70)
    assert struct.unpack("<i", msgs[2][2])[0] == 12
    assert struct.unpack("<f", msgs[3][2])[0] == np.float32(0.01)
    np.testing.assert_array_equal(np.frombuffer(msgs[4][2], "<f4"),
                                  np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 0.4], np.float32))
    np.testing.assert_array_equal(np.frombuffer(msgs[5][2], "<f4").reshape(n, 3), ref0)

    t = np.float32(0)
    for frame in range(3):
        c.text("run")
        op, fin, payload = c.recv()
        assert op == 2 and len(payload) == 4 * (3 * n + 6)
        floats = np.frombuffer(payload, "<f4")
        sim.take_step(5)
        for _ in range(5):
            t = np.float32(t + np.float32(0.0002))
        # the same engine from the same state: the same bits
        np.testing.assert_array_equal(floats[:3 * n].reshape(n, 3), sim.state()[:, :3].astype(np.float32))
        assert floats[3 * n] == t == np.float32(sim.time)
        assert np.all(np.isfinite(floats[3 * n + 1:])) and np.all(floats[3 * n + 2:] >= 0)
    # "initialize" again restarts from the initial condition
    c.text("initialize")
    msgs = [c.recv() for _ in range(6)]
    np.testing.assert_array_equal(np.frombuffer(msgs[5][2], "<f4").reshape(n, 3), ref0)
    c.close()
    sim.close()
