"""Reader for the .gdtraj files gd_run writes per steps_per_dump window.

A file holds a gd_traj_header (include/gd_remc.h) and n_rep x n_step
gd_step_record entries, replica-major: the records the reference dumps as
HDF5 per window (hdf5io.C:10-30, read back by analysis.C:13-104 and
readH5.py:6-33).
"""
import numpy as np

from .remc import STEP_RECORD_DTYPE

HEADER_DTYPE = np.dtype([("magic", "S8"), ("n_rep", "<i4"), ("n_step", "<i4"),
                         ("record_bytes", "<i4"), ("first_step", "<i4"), ("stride", "<i4"),
                         ("reserved", "<i4", (3,))])


def read_gdtraj(path):
    """(header dict, records [n_rep, n_step] of STEP_RECORD_DTYPE)."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER_DTYPE.itemsize:
        raise ValueError(f"{path}: shorter than the header")
    h = np.frombuffer(raw[:HEADER_DTYPE.itemsize], HEADER_DTYPE)[0]
    if h["magic"] != b"GDTRAJ01":
        raise ValueError(f"{path}: not a .gdtraj file")
    if h["record_bytes"] != STEP_RECORD_DTYPE.itemsize:
        raise ValueError(f"{path}: record size {h['record_bytes']}, expected "
                         f"{STEP_RECORD_DTYPE.itemsize}")
    n_rep, n_step = int(h["n_rep"]), int(h["n_step"])
    body = raw[HEADER_DTYPE.itemsize:]
    need = n_rep * n_step * STEP_RECORD_DTYPE.itemsize
    if len(body) != need:
        raise ValueError(f"{path}: {len(body)} record bytes, expected {need}")
    rec = np.frombuffer(body, STEP_RECORD_DTYPE).reshape(n_rep, n_step)
    header = {"n_rep": n_rep, "n_step": n_step, "first_step": int(h["first_step"]),
              "stride": int(h["stride"])}
    return header, rec


def write_gdtraj(path, first_step, stride, rec):
    """Write records [n_rep, n_step] in the same format (used to build test
    inputs and by tools that filter trajectories)."""
    rec = np.ascontiguousarray(rec, STEP_RECORD_DTYPE)
    h = np.zeros((), HEADER_DTYPE)
    h["magic"] = b"GDTRAJ01"
    h["n_rep"], h["n_step"] = rec.shape
    h["record_bytes"] = STEP_RECORD_DTYPE.itemsize
    h["first_step"], h["stride"] = first_step, stride
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(rec.tobytes())
