"""The DEM restart file <RUN_NAME>_DES.RES, written and read on the host.

Layout of WRITE_RES0_DES (write_res0_des.f:26-215) for a serial run
(bDIST_IO false, one process, no ghosts) of this library's physics (no
MPPIC, CGDEM, energy or species equations, glued or super-quadric
particles): an unformatted direct-access file of 512-byte records
(OPEN_N1, machine_mod.f).  A scalar fills the start of one record
(WRITE_RES_DES_0I / _0D); an array of N values fills ceil(N / 64) records
of doubles or ceil(N / 128) records of 4-byte integers, and at least one
record when N <= 64 or 128, even N = 0 (OUT_BIN_512 / OUT_BIN_512I).

  1 VERSION (1.2)        2 particles       3 ghosts (0)      4 collisions
  5 VTP_FINDEX           6 TECPLOT_FINDEX  7 DTSOLID
  DES_POS_NEW x, y, z; iGLOBAL_ID; PARTICLE_STATE; DES_VEL_NEW x, y, z;
  OMEGA_NEW x, y, z; DES_RADIUS; RO_SOL; PIJK(:,5)       (per particle)
  DES_USR_VAR_SIZE (scalar), then that many per-particle arrays;
  RESIDENCE_TIME                                          (per particle)
  two integer arrays per collision, then PFT_NEIGHBOR x, y, z
  DEM_BCMI (0: no mass inlets)

Per-particle arrays hold the existing particles in slot order
(DES_GATHER).  The collision arrays are read back by READ_PAR_COL
(read_res1_des_mod.f:508-570) as iPAR_COL(1,:), the owner's global id,
and iPAR_COL(2,:), the partner's; this module writes them so.  (The
reference's current writer puts NEIGHBORS and NEIGHBOR_INDEX there, and
its GLOBAL_TO_LOC_COL, :717-804, rebuilds no pair list from them; here the
pairs restart through mfixdem_set_pairs.)
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

RECL = 512
NWORDS_DP = 64
NWORDS_I = 128
VERSION = 1.2


@dataclasses.dataclass
class DesRes:
    version: float
    dtsolid: float
    pos: np.ndarray          # (3, n)
    gid: np.ndarray          # (n,) iGLOBAL_ID, 1-based
    state: np.ndarray        # (n,) PARTICLE_STATE
    vel: np.ndarray
    omega: np.ndarray
    radius: np.ndarray
    ro_sol: np.ndarray
    phase: np.ndarray        # (n,) PIJK(:,5), 1-based
    usr_var: np.ndarray      # (DES_USR_VAR_SIZE, n)
    residence_time: np.ndarray
    pair_owner: np.ndarray   # (m,) global id of the pair's owner (L)
    pair_partner: np.ndarray  # (m,) global id of the partner (I)
    pft: np.ndarray          # (m, 3) PFT_NEIGHBOR
    vtp_findex: int = 0
    tecplot_findex: int = 0

    @property
    def n(self) -> int:
        return int(self.radius.shape[0])


class _Records:
    def __init__(self):
        self.recs: list[bytes] = []

    def scalar_i(self, v: int):
        self.recs.append(struct.pack("<i", int(v)))

    def scalar_d(self, v: float):
        self.recs.append(struct.pack("<d", float(v)))

    def array(self, a, dtype, words):
        a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        if a.size <= words:
            self.recs.append(a.tobytes())
            return
        for s in range(0, a.size, words):
            self.recs.append(a[s:s + words].tobytes())

    def tobytes(self) -> bytes:
        return b"".join(r.ljust(RECL, b"\0") for r in self.recs)


def write_des_res(path: str, r: DesRes) -> None:
    n, m = r.n, int(r.pair_owner.shape[0])
    w = _Records()
    w.scalar_d(VERSION)
    w.scalar_i(n)
    w.scalar_i(0)
    w.scalar_i(m)
    w.scalar_i(r.vtp_findex)
    w.scalar_i(r.tecplot_findex)
    w.scalar_d(r.dtsolid)
    for d in range(3):
        w.array(r.pos[d], "<f8", NWORDS_DP)
    w.array(r.gid, "<i4", NWORDS_I)
    w.array(r.state, "<i4", NWORDS_I)
    for d in range(3):
        w.array(r.vel[d], "<f8", NWORDS_DP)
    for d in range(3):
        w.array(r.omega[d], "<f8", NWORDS_DP)
    w.array(r.radius, "<f8", NWORDS_DP)
    w.array(r.ro_sol, "<f8", NWORDS_DP)
    w.array(r.phase, "<i4", NWORDS_I)
    w.scalar_i(r.usr_var.shape[0])
    for k in range(r.usr_var.shape[0]):
        w.array(r.usr_var[k], "<f8", NWORDS_DP)
    w.array(r.residence_time, "<f8", NWORDS_DP)
    w.array(r.pair_owner, "<i4", NWORDS_I)
    w.array(r.pair_partner, "<i4", NWORDS_I)
    for d in range(3):
        w.array(r.pft[:, d], "<f8", NWORDS_DP)
    w.scalar_i(0)   # DEM_BCMI
    with open(path, "wb") as f:
        f.write(w.tobytes())


def read_des_res(path: str) -> DesRes:
    raw = open(path, "rb").read()
    if len(raw) % RECL:
        raise ValueError(f"{path}: {len(raw)} bytes is not a whole number of {RECL}-byte records")
    nrec = len(raw) // RECL
    pos_ = [0]

    def rec():
        if pos_[0] >= nrec:
            raise ValueError(f"{path}: ends before record {pos_[0] + 1}")
        b = raw[pos_[0] * RECL:(pos_[0] + 1) * RECL]
        pos_[0] += 1
        return b

    def si():
        return struct.unpack("<i", rec()[:4])[0]

    def sd():
        return struct.unpack("<d", rec()[:8])[0]

    def arr(count, dtype, words, size):
        out = []
        left = count
        if count <= words:
            out.append(np.frombuffer(rec()[:count * size], dtype=dtype))
        else:
            while left > 0:
                k = min(words, left)
                out.append(np.frombuffer(rec()[:k * size], dtype=dtype))
                left -= k
        return np.concatenate(out).astype(dtype.lstrip("<"))

    version = sd()
    if abs(version - VERSION) > 1e-6:
        raise ValueError(f"{path}: DES.RES version {version}, this reader takes {VERSION}")
    n, ghosts, m = si(), si(), si()
    if ghosts != 0:
        raise ValueError("distributed-memory restart files (ghosts) are not read")
    vtp, tec = si(), si()
    dt = sd()
    d = lambda: arr(n, "<f8", NWORDS_DP, 8)   # noqa: E731
    i4 = lambda c: arr(c, "<i4", NWORDS_I, 4)  # noqa: E731
    pos = np.stack([d() for _ in range(3)])
    gid = i4(n)
    state = i4(n)
    vel = np.stack([d() for _ in range(3)])
    omega = np.stack([d() for _ in range(3)])
    radius = d()
    ro_sol = d()
    phase = i4(n)
    nu = si()
    usr = np.stack([d() for _ in range(nu)]) if nu else np.zeros((0, n))
    res_t = d()
    owner = i4(m)
    partner = i4(m)
    pft = np.stack([arr(m, "<f8", NWORDS_DP, 8) for _ in range(3)], axis=1)
    bcmi = si()
    if bcmi != 0:
        raise ValueError("restart files with mass inlets (DEM_BCMI > 0) are not read")
    return DesRes(version, dt, pos, gid, state, vel, omega, radius, ro_sol, phase, usr, res_t,
                  owner, partner, pft, vtp, tec)


def from_system(sys_, radius, ro_sol, phase) -> DesRes:
    """Restart data of a DemSystem (or the oracle): the existing particles
    in global-id order (their ids kept), their input radius, density and
    phase, and every listed pair of two existing particles with its
    tangential history."""
    pos, vel, om, state = sys_.get_particles()
    keep = state != 0
    ids = np.nonzero(keep)[0]
    idx, nb, pft = sys_.get_pairs()
    ends = np.asarray(idx, dtype=np.int64)
    starts = np.concatenate([[1], ends[:-1]]) if ends.size else ends
    owner = np.repeat(np.arange(1, ends.size + 1), np.maximum(ends - starts, 0))
    live = keep[owner - 1] & keep[np.asarray(nb) - 1] if owner.size else np.zeros(0, bool)
    return DesRes(VERSION, float(sys_.derived().dtsolid), pos[:, keep], ids + 1,
                  state[keep].astype(np.int32), vel[:, keep], om[:, keep],
                  np.asarray(radius)[keep], np.asarray(ro_sol)[keep],
                  np.asarray(phase)[keep].astype(np.int32), np.zeros((0, ids.size)),
                  np.zeros(ids.size), owner[live].astype(np.int32),
                  np.asarray(nb)[live].astype(np.int32), np.asarray(pft)[live])
