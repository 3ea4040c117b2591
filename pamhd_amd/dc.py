"""PAMHD's .dc (version 4) result files for uniform grids.

Layout (tests/common.py:35-84, source/mhd/save.hpp): a header of
    u64 file_version (4), u64 sim_step,
    f64 sim_time, adiabatic_index, proton_mass, vacuum_permeability,
        particle_temp_nrj_ratio,
    u8 number of variables, u64 offset of each variable block,
    u64 endianness mark 0x1234567890abcdef,
    u64[3] level-0 cells, i32 max refinement level, u32 neighbourhood length,
    u8[3] periodicity, i32 geometry id (1, Cartesian),
    f64[3] grid start, f64[3] level-0 cell length,
    u64 number of cells, u64[n] cell ids (1 + i + nx (j + ny k) on level 0),
then one block per variable: an 8-character name and the values of every
cell in the order of the id list.  Variables written here:
    "mhd     "  rho, m (3), E, cell-centred B (3)        8 f64 per cell
    "faceB   "  face B of the -x, +x, -y, +y, -z, +z faces  6 f64 per cell
    "mhd info"  cell type (1 normal, 0 boundary, -1 not solved)  i32
"""
import struct

import numpy as np

ENDIAN_MARK = 0x1234567890abcdef
FORMATS = {"mhd     ": ("<f8", 8), "faceB   ": ("<f8", 6), "mhd info": ("<i4", 1),
           "divfaceB": ("<f8", 1), "Berror  ": ("<f8", 1), "timestep": ("<f8", 1),
           "rank    ": ("<i4", 1), "substep ": ("<i4", 1), "substmin": ("<i4", 1),
           "substmax": ("<i4", 1), "bgB     ": ("<f8", 18), "volE    ": ("<f8", 3),
           "volJ    ": ("<f8", 3), "ref lvls": ("<i4", 2), "nr ipart": ("<u8", 1)}


def cells_of_grid_state(state, cells):
    """(8, n) build state (face B on the negative faces) or (14, n) box state
    to the per-cell rows of "mhd" and "faceB"."""
    nx, ny, nz = cells
    n = nx * ny * nz
    if state.shape == (8, n):
        s = state.reshape(8, nz, ny, nx)
        plus = [np.roll(s[5], -1, 2), np.roll(s[6], -1, 1), np.roll(s[7], -1, 0)]
        faces = np.stack([s[5], plus[0], s[6], plus[1], s[7], plus[2]]).reshape(6, n)
        vb = 0.5 * (faces[0::2] + faces[1::2])
        mhd = np.concatenate([state[:5], vb])
    elif state.shape == (14, n):
        mhd = state[:8]
        faces = state[8:14]
    else:
        raise ValueError("state must be (8, n) or (14, n)")
    return mhd.T.copy(), faces.T.copy()


def write(path, state, cells, start, cell_length, periodic, sim_step=0, sim_time=0.0,
          adiabatic_index=5.0 / 3.0, proton_mass=1.0, vacuum_permeability=1.0,
          cell_type=None, particle_temp_nrj_ratio=0.0):
    """Writes a version-4 file of a uniform grid state."""
    nx, ny, nz = cells
    n = nx * ny * nz
    mhd, faces = cells_of_grid_state(np.asarray(state, dtype=np.float64), cells)
    blocks = [("mhd     ", mhd.astype("<f8")), ("faceB   ", faces.astype("<f8"))]
    if cell_type is not None:
        blocks.append(("mhd info", np.asarray(cell_type, dtype="<i4").reshape(n, 1)))
    head = struct.pack("<QQ5d", 4, sim_step, sim_time, adiabatic_index, proton_mass,
                       vacuum_permeability, particle_temp_nrj_ratio)
    nvar = len(blocks)
    rest = struct.pack("<Q3QiI3Bi3d3dQ", ENDIAN_MARK, nx, ny, nz, 0, 1,
                       *[1 if p else 0 for p in periodic], 1, *start, *cell_length, n)
    ids = (np.arange(n, dtype="<u8") + 1).tobytes()
    header_len = len(head) + 1 + 8 * nvar + len(rest) + len(ids)
    offsets, pos = [], header_len
    for name, data in blocks:
        offsets.append(pos)
        pos += 8 + data.nbytes
    with open(path, "wb") as f:
        f.write(head)
        f.write(struct.pack("<B", nvar))
        f.write(struct.pack("<%dQ" % nvar, *offsets))
        f.write(rest)
        f.write(ids)
        for name, data in blocks:
            f.write(name.encode("ascii"))
            f.write(data.tobytes())


def read(path):
    """Header dict, cell ids and {variable name: (n, k) array} of a v4 file."""
    with open(path, "rb") as f:
        raw = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, raw, off)
        off += struct.calcsize("<" + fmt)
        return v

    meta = {}
    meta["file_version"], meta["sim_step"] = take("QQ")
    if meta["file_version"] != 4:
        raise ValueError("unsupported file version %d" % meta["file_version"])
    (meta["sim_time"], meta["adiabatic_index"], meta["proton_mass"],
     meta["vacuum_permeability"], meta["particle_temp_nrj_ratio"]) = take("5d")
    nvar = take("B")[0]
    offsets = take("%dQ" % nvar)
    if take("Q")[0] != ENDIAN_MARK:
        raise ValueError("unsupported endianness")
    meta["ref_lvl_0_cells"] = take("3Q")
    meta["max_ref_lvl"] = take("i")[0]
    meta["neighborhood_length"] = take("I")[0]
    meta["periodicity"] = take("3B")
    meta["geometry_id"] = take("i")[0]
    meta["grid_start"] = take("3d")
    meta["lvl_0_cell_length"] = take("3d")
    ncell = take("Q")[0]
    ids = np.frombuffer(raw, dtype="<u8", count=ncell, offset=off).copy()
    data = {}
    for o in offsets:
        name = raw[o:o + 8].decode("ascii")
        if name not in FORMATS:
            continue
        dt, k = FORMATS[name]
        data[name] = np.frombuffer(raw, dtype=dt, count=k * ncell, offset=o + 8).reshape(ncell, k).copy()
    return meta, ids, data


def to_grid_order(ids, rows, cells):
    """Rows of a level-0 file in i-fastest grid order, (k, n)."""
    nx, ny, nz = cells
    out = np.zeros((rows.shape[1], nx * ny * nz), dtype=rows.dtype)
    out[:, ids.astype(np.int64) - 1] = rows.T
    return out
