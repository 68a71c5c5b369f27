"""ATTRACT receptor grid files: reader (both reference flavours), list padding, writer.

Reader behaviour follows read_grid.py:5-139 (CSR neighbour lists) and
read_grid2.py:16-161 (dense neighbour table); pad_grid follows read_grid.py:141-204.
The binary layout (arch 64, packed little-endian fields):

    bool is_torquegrid | short arch | double gridspacing | int gridextension |
    double plateaudis | double neighbourdis | bool alphabet[99] | float origin[3] |
    int dim[3] | int dim2[3] | int natoms | double pivot[3] |
    int nr_energrads | int shm_energrads | float energrads[nr][4] |
    int nr_neighbours | int shm_neighbours | {u8 type; pad3; u32 index}[nr] |
    long innergridsize | {u32 potential[100]; i32 neighbourlist; i16 nr_neighbours; pad2}[dz][dy][dx] |
    long biggridsize | u32 potential[dz2][dy2][dx2][100]

Inner/outer potential indices are 1-based into energrads (0 = none, outer grid only);
neighbourlist is a 1-based start into the neighbour array.
"""
import struct
from collections import namedtuple

import numpy as np

_HEADER = struct.Struct("<?hdidd99?3f3i3ii3d")
_INNER_DTYPE = np.dtype([("potential", "<u4", 100), ("neighbourlist", "<i4"), ("nr_neighbours", "<i2"),
                         ("_pad", "V2")])
_NB_DTYPE = np.dtype([("type", "u1"), ("_pad", "V3"), ("index", "<u4")])
assert _INNER_DTYPE.itemsize == 408 and _NB_DTYPE.itemsize == 8


class GridFormatError(AssertionError, ValueError):
    """Malformed grid data; an AssertionError like the reference reader's asserts."""


def _require(cond, msg):
    if not cond:
        raise GridFormatError(msg)


def _parse(grid_data):
    buf = memoryview(grid_data)
    _require(len(buf) >= _HEADER.size + 8, "grid data too short")
    h = _HEADER.unpack_from(buf, 0)
    pos = _HEADER.size
    is_torquegrid, arch = h[0], h[1]
    _require(not is_torquegrid, is_torquegrid)
    _require(arch == 64, arch)
    hdr = {
        "gridspacing": h[2],
        "gridextension": h[3],
        "plateaudis": h[4],
        "neighbourdis": h[5],
        "alphabet": np.array(h[6:105], bool),
        "origin": np.array(h[105:108]),
        "dim": np.array(h[108:111]),
        "dim2": np.array(h[111:114]),
        "natoms": h[114],
    }
    nr_energrads, shm_energrads = struct.unpack_from("<ii", buf, pos)
    pos += 8
    energrads = None
    if nr_energrads:
        _require(shm_energrads == -1, "Can't read grid from shared memory")
        energrads = np.frombuffer(buf, dtype="<f4", count=nr_energrads * 4, offset=pos).reshape(nr_energrads, 4)
        pos += energrads.nbytes
    nr_neighbours, shm_neighbours = struct.unpack_from("<ii", buf, pos)
    pos += 8
    _require(shm_neighbours == -1, "Can't read grid from shared memory")
    nb = np.frombuffer(buf, dtype=_NB_DTYPE, count=nr_neighbours, offset=pos)
    pos += nb.nbytes
    neighbours = np.ascontiguousarray(nb["index"]).astype(np.uint32)
    (innergridsize,) = struct.unpack_from("<q", buf, pos)
    pos += 8
    dx, dy, dz = (int(v) for v in hdr["dim"])
    _require(innergridsize == dx * dy * dz, (innergridsize, dx * dy * dz))
    inner = np.frombuffer(buf, dtype=_INNER_DTYPE, count=innergridsize, offset=pos)
    pos += inner.nbytes
    inner = np.ascontiguousarray(inner.reshape(dz, dy, dx).transpose(2, 1, 0))
    (biggridsize,) = struct.unpack_from("<q", buf, pos)
    pos += 8
    outer_ind = None
    if nr_energrads:
        dx2, dy2, dz2 = (int(v) for v in hdr["dim2"])
        _require(biggridsize == dx2 * dy2 * dz2, (biggridsize, dx2 * dy2 * dz2, (dx2, dy2, dz2)))
        outer_ind = np.frombuffer(buf, dtype="<u4", count=biggridsize * 100, offset=pos)
        pos += outer_ind.nbytes
        outer_ind = outer_ind.reshape(dz2, dy2, dx2, 100).transpose(2, 1, 0, 3)
    else:
        _require(biggridsize == 0, biggridsize)
    return hdr, energrads, neighbours, inner, outer_ind


def _potential_grids(hdr, energrads, inner, outer_ind, read_gradients):
    """Resolve per-type potential indices into dense grids (read_grid.py:79-135)."""
    alphabet = hdr["alphabet"]
    types = np.nonzero(alphabet)[0]
    src = energrads if read_gradients else np.ascontiguousarray(energrads[:, 0])
    n = len(energrads)
    pot_ind = inner["potential"]
    unused = np.nonzero(~alphabet)[0]
    if len(unused):
        bad = pot_ind[..., unused].max(axis=(0, 1, 2))
        for t, b in zip(unused, bad):
            _require(b == 0, (int(t), int(b)))
    inner_grid = np.empty((len(types),) + inner.shape + ((4,) if read_gradients else ()), np.float32)
    for k, t in enumerate(types):
        ind = pot_ind[..., t]
        _require(ind.min() >= 1 and ind.max() <= n, (int(t), int(ind.min()), int(ind.max()), n))
        inner_grid[k] = src[ind.astype(np.int64) - 1]
    outer_grid = np.zeros((len(types),) + outer_ind.shape[:3] + ((4,) if read_gradients else ()), np.float32)
    for t in unused:
        _require(outer_ind[..., t].max() == 0, (int(t), 0, int(outer_ind[..., t].max())))
    for k, t in enumerate(types):
        ind = outer_ind[..., t]
        _require(ind.max() <= n, (int(ind.max()), n))
        mask = ind > 0
        outer_grid[k][mask] = src[ind[mask].astype(np.int64) - 1]
    return inner_grid, outer_grid


def read_grid(grid_data, read_gradients=False):
    """Parse an ATTRACT grid into the reference's Grid namedtuple (read_grid.py:5-139).

    Fields, in order: gridspacing, gridextension, plateaudis, neighbourdis, alphabet,
    origin, dim, dim2, natoms, neighbours (uint32, 0-based receptor indices),
    [inner_potential_grid], neighbour_grid (dx,dy,dz,2 int32 = count, 1-based start),
    [outer_potential_grid].  Potential grids are (ntypes, dx, dy, dz[, 4]) float32 and
    are present only when the file holds energrads.
    """
    hdr, energrads, neighbours, inner, outer_ind = _parse(grid_data)
    d = dict(hdr)
    d["neighbours"] = neighbours
    if energrads is not None:
        inner_grid, outer_grid = _potential_grids(hdr, energrads, inner, outer_ind, read_gradients)
        d["inner_potential_grid"] = inner_grid
    neighbour_grid = np.empty(inner.shape + (2,), np.int32)
    neighbour_grid[..., 0] = inner["nr_neighbours"]
    neighbour_grid[..., 1] = inner["neighbourlist"]
    d["neighbour_grid"] = neighbour_grid
    if energrads is not None:
        d["outer_potential_grid"] = outer_grid
    return namedtuple("Grid", d.keys())(*d.values())


def read_grid2(grid_data, read_gradients=False):
    """Parse an ATTRACT grid with a dense neighbour table (read_grid2.py:16-161).

    Fields: ..., natoms, [inner_potential_grid], nr_neighbours (dx,dy,dz int16),
    max_nr_neighbours, neighbour_grid (dx,dy,dz,max_nr_neighbours uint16, padded with
    65535), [outer_potential_grid], neighbour_grid_ravel (None).
    Requires fewer than 2**16 receptor beads.
    """
    hdr, energrads, neighbours, inner, outer_ind = _parse(grid_data)
    if len(neighbours):
        _require(neighbours.max() < 2 ** 16, "Too many atoms")
    d = dict(hdr)
    if energrads is not None:
        inner_grid, outer_grid = _potential_grids(hdr, energrads, inner, outer_ind, read_gradients)
        d["inner_potential_grid"] = inner_grid
    nr = np.ascontiguousarray(inner["nr_neighbours"])
    d["nr_neighbours"] = nr
    kmax = int(nr.max()) if nr.size else 0
    d["max_nr_neighbours"] = nr.max() if nr.size else np.int16(0)
    table = np.full(inner.shape + (kmax,), 2 ** 16 - 1, np.uint16)
    if kmax:
        counts = np.maximum(nr.reshape(-1).astype(np.int64), 0)
        starts = inner["neighbourlist"].reshape(-1).astype(np.int64) - 1
        flat = table.reshape(-1, kmax)
        rows = np.repeat(np.arange(len(counts)), counts)
        cols = np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts)
        flat[rows, cols] = neighbours[np.repeat(starts, counts) + cols]
    d["neighbour_grid"] = table
    if energrads is not None:
        d["outer_potential_grid"] = outer_grid
    return namedtuple("Grid", tuple(d.keys()) + ("neighbour_grid_ravel",))(*d.values(), neighbour_grid_ravel=None)


def pad_grid(grid, padding):
    """Re-lay the CSR neighbour lists with extra 0xFFFFFFFF slots (read_grid.py:141-204).

    A list of length N gets padding[N-1] extra slots; an integer padding is used for
    every length; a short padding list repeats its last element.  Starts stay 1-based,
    0 for an empty list.
    """
    ng = np.array(grid.neighbour_grid)
    g = ng.reshape(-1, 2)
    lengths = g[:, 0].astype(np.int64)
    max_contacts = int(lengths.max())
    if np.ndim(padding) == 0:
        pad = np.full(max_contacts, int(padding), np.int64)
    else:
        padding = list(padding)
        pad = np.empty(max_contacts, np.uint16)
        lp = min(max_contacts, len(padding))
        pad[:lp] = padding[:lp]
        pad[len(padding):max_contacts] = padding[-1]
        pad = pad.astype(np.int64)
    padded = np.where(lengths > 0, lengths + pad[lengths - 1], 0)
    pos = np.concatenate(([0], np.cumsum(padded)[:-1]))
    old = g[:, 1].astype(np.int64) - 1
    nb2 = np.full(int(padded.sum()), np.iinfo(np.uint32).max, np.uint32)
    nz = lengths > 0
    ln = lengths[nz]
    within = np.arange(ln.sum()) - np.repeat(np.cumsum(ln) - ln, ln)
    nb2[np.repeat(pos[nz], ln) + within] = np.asarray(grid.neighbours)[np.repeat(old[nz], ln) + within]
    g[:, 1] = np.where(nz, pos + 1, 0)
    return grid._replace(neighbours=nb2, neighbour_grid=ng)


def write_grid(*, gridspacing, gridextension, plateaudis, neighbourdis, alphabet, origin, dim, dim2, natoms,
               inner_potential_grid=None, outer_potential_grid=None, nb_counts, neighbours, nb_types=None,
               outer_mask=None):
    """Serialise a grid in the ATTRACT binary format read by read_grid.

    inner/outer_potential_grid: (ntypes, d..., 4) float32 [E, gx, gy, gz] for the
    alphabet types in order (None for a neighbour-only .nbgrid).  nb_counts:
    (dx,dy,dz) list lengths; neighbours: flat 0-based receptor indices, lists in
    x-major voxel order.  outer_mask (optional, dx2,dy2,dz2 bool) marks outer points
    that get no potential (index 0).  Every (type, point) gets its own energrad record.
    """
    alphabet = np.asarray(alphabet, bool)
    _require(alphabet.shape == (99,), "alphabet must have 99 entries")
    dim = [int(v) for v in dim]
    dim2 = [int(v) for v in dim2]
    types = np.nonzero(alphabet)[0]
    nt = len(types)
    out = bytearray()
    out += _HEADER.pack(False, 64, float(gridspacing), int(gridextension), float(plateaudis), float(neighbourdis),
                        *[bool(a) for a in alphabet], *[float(o) for o in origin], *dim, *dim2, int(natoms),
                        0.0, 0.0, 0.0)
    has_pot = inner_potential_grid is not None
    nin = dim[0] * dim[1] * dim[2]
    nout = dim2[0] * dim2[1] * dim2[2]
    if has_pot:
        ig = np.asarray(inner_potential_grid, np.float32)
        og = np.asarray(outer_potential_grid, np.float32)
        _require(ig.shape == (nt, *dim, 4) and og.shape == (nt, *dim2, 4), "potential grid shape mismatch")
        # energrad records: inner (type, z, y, x) then outer (type, z, y, x)
        e_in = ig.transpose(0, 3, 2, 1, 4).reshape(-1, 4)
        e_out = og.transpose(0, 3, 2, 1, 4).reshape(-1, 4)
        energrads = np.concatenate([e_in, e_out]).astype("<f4")
        out += struct.pack("<ii", len(energrads), -1)
        out += energrads.tobytes()
    else:
        out += struct.pack("<ii", 0, -1)
    neighbours = np.asarray(neighbours, np.uint32)
    rec = np.zeros(len(neighbours), _NB_DTYPE)
    rec["index"] = neighbours
    if nb_types is not None:
        rec["type"] = np.asarray(nb_types, np.uint8)
    out += struct.pack("<ii", len(neighbours), -1)
    out += rec.tobytes()
    counts = np.asarray(nb_counts).reshape(dim)
    inner = np.zeros((dim[2], dim[1], dim[0]), _INNER_DTYPE)
    c_zyx = counts.transpose(2, 1, 0)
    starts = np.concatenate(([0], np.cumsum(counts.reshape(-1))[:-1])).reshape(dim).transpose(2, 1, 0)
    inner["nr_neighbours"] = c_zyx
    inner["neighbourlist"] = np.where(c_zyx > 0, starts + 1, 0)
    if has_pot:
        pot = np.zeros((dim[2], dim[1], dim[0], 100), np.uint32)
        flat = np.arange(nin, dtype=np.uint32).reshape(dim[2], dim[1], dim[0])
        for k, t in enumerate(types):
            pot[..., t] = 1 + k * nin + flat
        inner["potential"] = pot
    out += struct.pack("<q", nin)
    out += inner.tobytes()
    if has_pot:
        out += struct.pack("<q", nout)
        pot = np.zeros((dim2[2], dim2[1], dim2[0], 100), np.uint32)
        flat = np.arange(nout, dtype=np.uint32).reshape(dim2[2], dim2[1], dim2[0])
        base = nt * nin
        for k, t in enumerate(types):
            pot[..., t] = 1 + base + k * nout + flat
        if outer_mask is not None:
            pot[np.asarray(outer_mask).transpose(2, 1, 0)] = 0
        out += pot.tobytes()
    else:
        out += struct.pack("<q", 0)
    return bytes(out)
