"""Navier-Stokes and Maxwell restart files of SFEMaNS (``suite_ns*``, ``suite_maxwell*``).

The files are Fortran sequential unformatted files: each record is framed by
its length in bytes as a 4-byte little-endian integer, before and after.
Layout (SOURCES/INITIALISATION/restart.F90, write_restart_ns :6-104,
read_restart_ns :106-337):

* header record: ``time, nb_procs_S, nb_procs_F, m_max_c [, dt]`` (per
  meridian sub-domain files) or ``time, np_v, np_p, nb_procs_F, m_max_c
  [, dt]`` (``mono``: one file for the whole mesh) (:64-75, :193-205);
* then, for every mode written by every Fourier rank in rank order: the mode
  number, ``un(:,:,i)`` (np_v, 6), ``un_m1``, ``pn(:,:,i)`` (np_p, 2),
  ``pn_m1``, ``incpn``, ``incpn_m1`` and, with level sets,
  ``level_set(:,:,:,i)`` (n_int, np_v, 2), ``level_set_m1`` and ``max_vel``
  (:82-94).

Maxwell restarts (``suite_maxwell*``, write_restart_maxwell :1019-1058,
read_restart_maxwell :1064-1275) have the same header; per mode the mode
number, Hn, Hn1, Bn, Bn1 (np_H, 6) and phin, phin1 (np_phi, 2), each
replaced by a record holding the integer 1 when its mesh is empty.

The readers return the fields for the modes the caller asks for, with the
reference's rules: a mode absent from the file is filled with ``val_init``
or 0 (:284-306); with ``interpol`` the i-th mode of the rank's share is
taken whatever its number and ``list_mode`` is updated (:244-255); with
``dt`` different from the file's (relative 1e-4), the ``_m1`` fields are
extrapolated to the new step (:316-335).  Everything here is host IO: the
fields come back as numpy arrays in the reference's layout, ready for
``fortran_tensor`` and the GPU entry points.
"""
from __future__ import annotations

import struct

import numpy as np

__all__ = ["suite_ns_name", "suite_maxwell_name", "read_records", "write_records", "read_restart_ns",
           "write_restart_ns", "read_restart_maxwell", "write_restart_maxwell"]


def suite_ns_name(filename: str, rang_S: int = 0, mono: bool = False, it: int | None = None) -> str:
    """File name of the NS restart (restart.F90:157-176): suite_ns[_S<rank>][_I<it>].<filename>,
    three-digit zero-padded numbers."""
    parts = "suite_ns"
    if not mono:
        parts += f"_S{rang_S:03d}"
    if it is not None:
        parts += f"_I{it:03d}"
    return f"{parts}.{filename}"


def suite_maxwell_name(filename: str, rang_S: int = 0, mono: bool = False, it: int | None = None) -> str:
    """File name of the Maxwell restart: suite_maxwell[_S<rank>][_I<it>].<filename>."""
    return "suite_maxwell" + suite_ns_name(filename, rang_S, mono, it)[len("suite_ns"):]


def read_records(path: str) -> list[bytes]:
    """All records of a Fortran sequential unformatted file (4-byte markers)."""
    data = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        (n,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data):
            raise ValueError(f"{path}: bad record length {n} at byte {pos}")
        (n2,) = struct.unpack_from("<i", data, end)
        if n2 != n:
            raise ValueError(f"{path}: record markers {n} and {n2} differ at byte {pos}")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def write_records(path: str, records, append: bool = False) -> None:
    with open(path, "ab" if append else "wb") as f:
        for r in records:
            f.write(struct.pack("<i", len(r)))
            f.write(r)
            f.write(struct.pack("<i", len(r)))


def _farray(rec: bytes, shape) -> np.ndarray:
    a = np.frombuffer(rec, dtype="<f8")
    if a.size != int(np.prod(shape)):
        raise ValueError(f"record of {a.size} doubles, expected {shape}")
    return a.reshape(shape, order="F").astype(np.float64)


def _header(rec: bytes, mono: bool, with_dt: bool):
    if mono:
        fmt = "<diiii" + ("d" if with_dt else "")
        v = struct.unpack(fmt, rec[:struct.calcsize(fmt)])
        return {"time": v[0], "np_v": v[1], "np_p": v[2], "nb_procs_S": -1, "nb_procs_F": v[3],
                "m_max_c": v[4], "dt": v[5] if with_dt else None}
    fmt = "<diii" + ("d" if with_dt else "")
    v = struct.unpack(fmt, rec[:struct.calcsize(fmt)])
    return {"time": v[0], "nb_procs_S": v[1], "nb_procs_F": v[2], "m_max_c": v[3],
            "dt": v[4] if with_dt else None}


def write_restart_ns(path, time, list_mode, un, un_m1, pn, pn_m1, incpn, incpn_m1, level_set=None,
                     level_set_m1=None, max_vel=0.0, dt=None, mono=False, nb_procs_S=1, nb_procs_F=1,
                     append=False):
    """Write (one Fourier rank's share of) a restart file as write_restart_ns
    does (restart.F90:59-97).  un (np_v, 6, m), pn/incpn (np_p, 2, m),
    level sets (n_int, np_v, 2, m).  The header is written unless ``append``
    (ranks after the first append their modes)."""
    un = np.asarray(un, dtype=np.float64)
    pn = np.asarray(pn, dtype=np.float64)
    recs = []
    if not append:
        if mono:
            h = struct.pack("<diiii", float(time), un.shape[0], pn.shape[0], nb_procs_F, len(list_mode))
        else:
            h = struct.pack("<diii", float(time), nb_procs_S, nb_procs_F, len(list_mode))
        if dt is not None:
            h += struct.pack("<d", float(dt))
        recs.append(h)
    with_ls = level_set is not None and level_set_m1 is not None
    for i, mode in enumerate(list_mode):
        recs.append(struct.pack("<i", int(mode)))
        for a in (un, un_m1, pn, pn_m1, incpn, incpn_m1):
            recs.append(np.asarray(a, dtype="<f8")[..., i].tobytes(order="F"))
        if with_ls:
            recs.append(np.asarray(level_set, dtype="<f8")[..., i].tobytes(order="F"))
            recs.append(np.asarray(level_set_m1, dtype="<f8")[..., i].tobytes(order="F"))
            recs.append(struct.pack("<d", float(max_vel)))
    write_records(path, recs, append=append)


def read_restart_ns(path, list_mode, np_v, np_p, n_int=None, mono=False, val_init=None, interpol=False,
                    dt=None, rang_F=0, nb_procs_S=None):
    """Read the modes ``list_mode`` (this rank's share) as read_restart_ns does.

    n_int: number of level sets stored per mode (None: no level sets).
    nb_procs_S: when given and the file is per sub-domain, it must match the
    file's (restart.F90:207-210).  Returns a dict with time, list_mode (as
    possibly updated by ``interpol``), un, un_m1 (np_v, 6, m), pn, pn_m1,
    incpn, incpn_m1 (np_p, 2, m), and with level sets level_set,
    level_set_m1 (n_int, np_v, 2, m) and max_vel; plus the header values."""
    recs = read_records(path)
    if not recs:
        raise ValueError(f"{path}: empty restart file")
    hd = _header(recs[0], mono, dt is not None)
    if nb_procs_S is not None and not mono and hd["nb_procs_S"] != nb_procs_S:
        raise ValueError("BUG in read_restart: nb_procs_Sr /= nb_procs_S")
    list_mode = [int(m) for m in list_mode]
    M = len(list_mode)
    nl = 6 + (3 if n_int is not None else 0)
    nb_mode_r = hd["nb_procs_F"] * hd["m_max_c"]
    if len(recs) < 1 + nb_mode_r * (1 + nl):
        raise ValueError(f"{path}: {len(recs)} records, expected {1 + nb_mode_r * (1 + nl)}")
    shapes = [(np_v, 6), (np_v, 6), (np_p, 2), (np_p, 2), (np_p, 2), (np_p, 2)]
    names = ["un", "un_m1", "pn", "pn_m1", "incpn", "incpn_m1"]
    out = {n: np.zeros(s + (M,), order="F") for n, s in zip(names, shapes)}
    if n_int is not None:
        out["level_set"] = np.zeros((n_int, np_v, 2, M), order="F")
        out["level_set_m1"] = np.zeros((n_int, np_v, 2, M), order="F")
    max_vel = 0.0
    for i in range(M):
        found = False
        for j in range(nb_mode_r):
            base = 1 + j * (1 + nl)
            (mode,) = struct.unpack("<i", recs[base][:4])
            if interpol:
                if j != rang_F * M + i:
                    continue
                list_mode[i] = mode
            if mode != list_mode[i]:
                continue
            for k, (n, s) in enumerate(zip(names, shapes)):
                out[n][..., i] = _farray(recs[base + 1 + k], s)
            if n_int is not None:
                out["level_set"][..., i] = _farray(recs[base + 7], (n_int, np_v, 2))
                out["level_set_m1"][..., i] = _farray(recs[base + 8], (n_int, np_v, 2))
                (max_vel,) = struct.unpack("<d", recs[base + 9][:8])
            found = True
            break
        if not found:
            fill = 0.0 if val_init is None else float(val_init)
            for n in out:
                out[n][..., i] = fill
            if val_init is not None and n_int is not None:
                max_vel = fill
    if dt is not None and abs((dt - hd["dt"]) / dt) > 1e-4:
        r = dt / hd["dt"]
        for n in ("un", "pn", "incpn") + (("level_set",) if n_int is not None else ()):
            out[n + "_m1"] = r * out[n + "_m1"] + (1.0 - r) * out[n]
    res = dict(hd)
    res.update(out)
    res["list_mode"] = list_mode
    if n_int is not None:
        res["max_vel"] = max_vel
    return res


def write_restart_maxwell(path, time, list_mode, Hn, Hn1, Bn, Bn1, phin, phin1, dt=None, mono=False, nb_procs_S=1,
                          nb_procs_F=1, append=False, H_empty=False, phi_empty=False):
    """Write a Maxwell restart as write_restart_maxwell does (restart.F90:1019-1058):
    per mode the mode number, Hn, Hn1, Bn, Bn1 (np_H, 6) or four records
    holding the integer 1 when the H mesh is empty, then phin, phin1
    (np_phi, 2) or two such records when the phi mesh is empty."""
    Hn = np.asarray(Hn, dtype=np.float64)
    phin = np.asarray(phin, dtype=np.float64)
    one = struct.pack("<i", 1)
    recs = []
    if not append:
        if mono:
            h = struct.pack("<diiii", float(time), 0 if H_empty else Hn.shape[0], 0 if phi_empty else phin.shape[0],
                            nb_procs_F, len(list_mode))
        else:
            h = struct.pack("<diii", float(time), nb_procs_S, nb_procs_F, len(list_mode))
        if dt is not None:
            h += struct.pack("<d", float(dt))
        recs.append(h)
    for i, mode in enumerate(list_mode):
        recs.append(struct.pack("<i", int(mode)))
        for a in (Hn, Hn1, Bn, Bn1):
            recs.append(one if H_empty else np.asarray(a, dtype="<f8")[..., i].tobytes(order="F"))
        for a in (phin, phin1):
            recs.append(one if phi_empty else np.asarray(a, dtype="<f8")[..., i].tobytes(order="F"))
    write_records(path, recs, append=append)


def read_restart_maxwell(path, list_mode, np_H, np_phi, mono=False, val_init=None, interpol=False, dt=None,
                         rang_F=0, nb_procs_S=None):
    """Read the modes ``list_mode`` of a Maxwell restart as read_restart_maxwell
    does (restart.F90:1064-1275).  np_H = 0 (np_phi = 0) stands for an empty
    H (phi) mesh, whose records are skipped.  Returns Hn, Hn1, Bn, Bn1
    (np_H, 6, m), phin, phin1 (np_phi, 2, m), list_mode and the header
    values; missing modes, val_init, interpol and dt as for read_restart_ns."""
    recs = read_records(path)
    if not recs:
        raise ValueError(f"{path}: empty restart file")
    hd = _header(recs[0], mono, dt is not None)
    if nb_procs_S is not None and not mono and hd["nb_procs_S"] != nb_procs_S:
        raise ValueError("BUG in read_restart: nb_procs_Sr /= nb_procs_S")
    list_mode = [int(m) for m in list_mode]
    M = len(list_mode)
    nb_mode_r = hd["nb_procs_F"] * hd["m_max_c"]
    if len(recs) < 1 + nb_mode_r * 7:
        raise ValueError(f"{path}: {len(recs)} records, expected {1 + nb_mode_r * 7}")
    names = [("Hn", (np_H, 6)), ("Hn1", (np_H, 6)), ("Bn", (np_H, 6)), ("Bn1", (np_H, 6)),
             ("phin", (np_phi, 2)), ("phin1", (np_phi, 2))]
    out = {n: np.zeros(sh + (M,), order="F") for n, sh in names}
    for i in range(M):
        found = False
        for j in range(nb_mode_r):
            base = 1 + j * 7
            (mode,) = struct.unpack("<i", recs[base][:4])
            if interpol:
                if j != rang_F * M + i:
                    continue
                list_mode[i] = mode
            if mode != list_mode[i]:
                continue
            for k, (n, sh) in enumerate(names):
                if sh[0] > 0:
                    out[n][..., i] = _farray(recs[base + 1 + k], sh)
            found = True
            break
        if not found:
            fill = 0.0 if val_init is None else float(val_init)
            for n in out:
                out[n][..., i] = fill
    if dt is not None and abs((dt - hd["dt"]) / dt) > 1e-4:
        r = dt / hd["dt"]
        for a, b in (("Hn1", "Hn"), ("Bn1", "Bn"), ("phin1", "phin")):
            out[a] = r * out[a] + (1.0 - r) * out[b]
    res = dict(hd)
    res.update(out)
    res["list_mode"] = list_mode
    return res
