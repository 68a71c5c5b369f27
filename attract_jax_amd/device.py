"""Device-resident receptor grid, pair table, receptor and ligands.

Torch tensors only hold device memory here; every computation runs in the HIP
kernels of libattract_amd.so.  Layouts in HBM:

* potentials: ATJ_LAYOUT_CORNER packs the 8 corners of each grid cell into one
  128-byte record (one cache line per bead per trilinear gather, 8x the plain
  footprint) or ATJ_LAYOUT_PLAIN keeps the reference (type, x, y, z, 4) layout;
* neighbour lists: per inner voxel (count, start) int32 pairs; the lists are
  uint16 receptor indices, each non-empty list starting on a 64-byte boundary and
  padded with 0xFFFF to a multiple of 32 entries (16-byte loads of 8 neighbours);
* receptor beads: float4 (x, y, z, pair-table row as int32 bits);
* ligand beads: float4 (x, y, z, grid type | pair-table column << 16 as int32 bits).
"""
import copy
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

SEG = 16  # lists are padded with 0xFFFF to a multiple of SEG entries (the longest segment read)
# Non-empty lists start on a 64-byte boundary (32 entries), so a list of up to 32 entries
# is one memory line and a longer one the fewest lines; the head of the list array is
# one such line of padding (read by idle lanes).
LINE = 32


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(device):
    return torch.device(device if device is not None else "cuda")


def _upload(a, device):
    """A host array on the device, copied on the current stream from pinned memory: the
    host does not wait for the copy, nor for work queued before it on the stream (a copy
    from pageable memory would wait for both)."""
    d = _dev(device)
    t = torch.from_numpy(np.ascontiguousarray(a))
    if d.type != "cuda":
        return t.to(d)
    return t.pin_memory().to(d, non_blocking=True)


def pack_beads(xyz, tag):
    """float4 records (x, y, z, tag as int32 bits) as a float32 (n, 4) numpy array."""
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    out = np.empty((len(xyz), 4), np.float32)
    out[:, :3] = xyz
    out[:, 3] = np.asarray(tag, np.int32).reshape(-1).view(np.float32)
    return out


@dataclass
class Receptor:
    beads: torch.Tensor  # (nrec, 4) float32
    nrec: int
    max_type: int = None  # largest pair-table row, when known on the host

    @classmethod
    def create(cls, coor_rec, rec_atomtypes, device=None):
        """coor_rec (nrec, 3); rec_atomtypes = pair-table row of every bead (compacted types)."""
        t = np.asarray(rec_atomtypes, np.int64).reshape(-1)
        if t.min(initial=0) < 0 or t.max(initial=0) > 0x7FFFFFFF:
            raise ValueError("receptor types out of range")
        b = pack_beads(coor_rec, t)
        return cls(_upload(b, device), len(b), int(t.max(initial=0)))

    def check_types(self, ff):
        """ValueError unless every bead's type is a row of the pair table ff."""
        if self.max_type is not None and self.max_type >= ff.nrt:
            raise ValueError(f"receptor type {self.max_type} beyond the {ff.nrt} rows of the pair table")

    def struct(self):
        return _lib.AtjReceptor(self.beads.data_ptr() if self.nrec else 0, self.nrec)


@dataclass
class Ligand:
    beads: torch.Tensor  # (nbeads, 4) float32
    nbeads: int
    coor: np.ndarray = field(repr=False, default=None)
    max_grid_type: int = None  # largest potential-grid index and pair-table column, when known
    max_ff_type: int = None

    @classmethod
    def create(cls, coor_lig, lig_atomtype_pos, lig_atomtypes, device=None, stream=None):
        """coor_lig (nlig, 3) local coordinates; lig_atomtype_pos = potential-grid index of
        every bead; lig_atomtypes = pair-table column of every bead.  With a stream, the beads
        are uploaded on it from pinned memory without waiting (a copy from pageable memory on
        the default stream would wait for work on the other streams)."""
        gt = np.asarray(lig_atomtype_pos, np.int64)
        lt = np.asarray(lig_atomtypes, np.int64)
        if np.asarray(coor_lig).size == 0:
            raise ValueError("a ligand needs at least one bead")
        if gt.min(initial=0) < 0 or gt.max(initial=0) > 0xFFFF or lt.min(initial=0) < 0 or lt.max(initial=0) > 0x7FFF:
            raise ValueError("ligand types out of range")
        b = pack_beads(coor_lig, gt | (lt << 16))
        if stream is None:
            beads = torch.from_numpy(b).to(_dev(device))
        else:
            host = torch.from_numpy(b).pin_memory()
            with torch.cuda.stream(stream):
                beads = torch.empty(host.shape, dtype=host.dtype, device=_dev(device))
                beads.copy_(host, non_blocking=True)
        return cls(beads, len(b), np.asarray(coor_lig, np.float32).reshape(-1, 3), int(gt.max(initial=0)),
                   int(lt.max(initial=0)))

    def check_types(self, grid, ff):
        """ValueError unless every bead's grid type is a potential grid of `grid` and its
        force-field type a column of the pair table ff."""
        if self.max_grid_type is not None and grid.has_potentials and self.max_grid_type >= grid.ntypes:
            raise ValueError(f"ligand grid type {self.max_grid_type} beyond the {grid.ntypes} potential grids")
        if self.max_ff_type is not None and self.max_ff_type >= ff.nlt:
            raise ValueError(f"ligand type {self.max_ff_type} beyond the {ff.nlt} columns of the pair table")

    def struct(self):
        return _lib.AtjLigand(self.beads.data_ptr(), self.nbeads, 0)


@dataclass
class ForceField:
    """Compacted ATTRACT pair table (test-grid-energrad.py:67-79) on the device."""
    rc: torch.Tensor
    ac: torch.Tensor
    ivor: torch.Tensor
    emin: torch.Tensor
    rmin2: torch.Tensor

    @property
    def nrt(self):
        return self.rc.shape[0]

    @property
    def nlt(self):
        return self.rc.shape[1]

    @classmethod
    def create(cls, rc, ac, ivor, emin=None, rmin2=None, device=None):
        """emin / rmin2 default to the reference's float64 derivation from rc, ac."""
        rc64 = np.asarray(rc, np.float64)
        ac64 = np.asarray(ac, np.float64)
        if emin is None:
            emin = -27.0 * ac64 ** 4 / (256.0 * rc64 ** 3)
        if rmin2 is None:
            rmin2 = 4.0 * rc64 / (3.0 * ac64)
        if not np.all(np.abs(np.asarray(ivor)) == 1):
            raise ValueError("ivor must be +1 or -1 (ATTRACT ipon)")
        f = lambda a: _upload(np.asarray(np.asarray(a, np.float64), np.float32), device)
        return cls(f(rc), f(ac), f(ivor), f(emin), f(rmin2))

    @classmethod
    def from_tensors(cls, rc, ac, ivor, emin, rmin2):
        c = lambda t: t.detach().to(torch.float32).contiguous()
        return cls(c(rc), c(ac), c(ivor), c(emin), c(rmin2))

    def struct(self):
        return _lib.AtjForceField(self.rc.data_ptr(), self.ac.data_ptr(), self.ivor.data_ptr(),
                                  self.emin.data_ptr(), self.rmin2.data_ptr(), self.nrt, self.nlt)


def lists_from_csr(counts, starts0, flat):
    """Device neighbour-list format from CSR lists (counts, 0-based starts, flat indices):
    returns (nbgrid (V,2) int32, nblist uint16) with 0xFFFF padding to a multiple of 32
    entries, non-empty lists starting at multiples of 32 entries (>= 32) and an
    all-padding head of 32 entries."""
    counts = np.asarray(counts, np.int64).reshape(-1)
    counts = np.maximum(counts, 0)
    starts0 = np.asarray(starts0, np.int64).reshape(-1)
    flat = np.asarray(flat)
    slot = np.where(counts > 0, (counts + LINE - 1) // LINE * LINE, 0)
    new_starts = LINE + np.concatenate(([0], np.cumsum(slot)[:-1]))
    total = LINE + int(slot.sum())
    nblist = np.full(total, 0xFFFF, np.uint16)
    n = int(counts.sum())
    if n:
        within = np.arange(n) - np.repeat(np.cumsum(counts) - counts, counts)
        entries = flat[np.repeat(starts0, counts) + within]
        # only the entries the counts cover (a dense table's padding and pad_grid's sentinels
        # lie beyond them)
        if int(entries.max()) >= 0xFFFF:
            raise ValueError("neighbour lists need receptor indices < 65535")
        nblist[np.repeat(new_starts, counts) + within] = entries.astype(np.uint16)
    nbgrid = np.stack([counts, new_starts], axis=1).astype(np.int32)
    return nbgrid, nblist


def lists_from_grid(grid):
    """Device list format from a reference Grid: CSR (read_grid / pad_grid) or dense (read_grid2)."""
    if hasattr(grid, "nr_neighbours"):  # read_grid2 dense table
        counts = np.asarray(grid.nr_neighbours).reshape(-1).astype(np.int64)
        table = np.asarray(grid.neighbour_grid)
        k = table.shape[-1]
        flat = table.reshape(-1)
        starts0 = np.arange(len(counts), dtype=np.int64) * k
        return lists_from_csr(counts, starts0, flat)
    ng = np.asarray(grid.neighbour_grid).reshape(-1, 2)
    counts = ng[:, 0].astype(np.int64)
    starts0 = ng[:, 1].astype(np.int64) - 1
    starts0[counts <= 0] = 0
    return lists_from_csr(counts, starts0, np.asarray(grid.neighbours))


class DeviceGrid:
    """A receptor grid resident in HBM, in the kernels' layout."""

    def __init__(self, *, origin, spacing, dim, dim2, gridextension, plateaudis, neighbourdis, ntypes,
                 inner_plain=None, outer_plain=None, nbgrid=None, nblist=None, layout=_lib.ATJ_LAYOUT_CORNER,
                 device=None, keep_plain=False):
        self.device = _dev(device)
        self.origin = np.asarray(origin, np.float32)
        self.spacing = float(np.float32(spacing))
        self.dim = [int(v) for v in dim]
        self.dim2 = [int(v) for v in dim2]
        self.gridextension = float(gridextension)
        self.plateaudis = float(plateaudis)
        self.plateaudis2 = float(np.float32(float(plateaudis) ** 2))
        self.neighbourdis = float(neighbourdis)
        self.ntypes = int(ntypes)
        self.layout = layout
        self.inner_plain = inner_plain
        self.outer_plain = outer_plain
        self.nbgrid = nbgrid
        self.nblist = nblist
        self.nshell = 0
        self.nbshell = None
        self.inner = self._relayout(inner_plain, self.dim)
        self.outer = self._relayout(outer_plain, self.dim2) if outer_plain is not None else None
        if not keep_plain and layout != _lib.ATJ_LAYOUT_PLAIN:
            self.inner_plain = None
            self.outer_plain = None

    def _relayout(self, plain, dim):
        if plain is None:
            return None
        if self.layout == _lib.ATJ_LAYOUT_PLAIN:
            return plain
        packed = torch.empty((self.ntypes, dim[0] - 1, dim[1] - 1, dim[2] - 1, 8, 4), dtype=torch.float32,
                             device=self.device)
        lib = _lib.load()
        d = (ctypes.c_int32 * 3)(*dim)
        _lib.check(lib.atj_grid_pack_corners(_ptr(plain), self.ntypes, ctypes.byref(d), _ptr(packed),
                                             _stream(self.device)))
        return packed

    def with_layout(self, layout):
        """This grid in another potential layout, sharing the neighbour lists and their shell
        order (a view: the lists are not copied).  The plain potentials must be kept
        (keep_plain=True) unless this grid is plain already."""
        if layout == self.layout:
            return self
        if self.inner is not None and self.inner_plain is None:
            raise ValueError("the plain potentials were not kept (keep_plain=True)")
        v = copy.copy(self)
        v.layout = layout
        if layout == _lib.ATJ_LAYOUT_PLAIN:
            v.inner, v.outer = self.inner_plain, self.outer_plain
        else:
            v.inner = v._relayout(self.inner_plain, self.dim)
            v.outer = v._relayout(self.outer_plain, self.dim2) if self.outer_plain is not None else None
        return v

    @property
    def has_potentials(self):
        return self.inner is not None

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in (self.inner, self.outer, self.nbgrid, self.nblist,
                                                          self.nbshell,
                                                          self.inner_plain, self.outer_plain) if t is not None)

    def struct(self, whole_lists=False):
        """AtjGrid of this grid; with whole_lists the kernels read every list entry even when
        shell_lists made shell counts."""
        g = _lib.AtjGrid()
        g.origin[:] = [float(v) for v in self.origin]
        g.spacing = self.spacing
        g.dim[:] = self.dim
        g.dim2[:] = self.dim2
        g.gridextension = self.gridextension
        g.plateaudis = self.plateaudis
        g.plateaudis2 = self.plateaudis2
        g.ntypes = self.ntypes
        g.layout = self.layout
        g.inner = self.inner.data_ptr() if self.inner is not None else 0
        g.outer = self.outer.data_ptr() if self.outer is not None else 0
        g.nbgrid = self.nbgrid.data_ptr()
        g.nblist = self.nblist.data_ptr()
        if self.nshell > 0 and not whole_lists:
            g.nshell = self.nshell
            g.nbshell = self.nbshell.data_ptr()
        return g

    def shell_lists(self, receptor, nshell):
        """Reorder every list into distance shells around its grid point (nshell 1..6; 0
        keeps the current order and makes beads read whole lists again).

        A bead in the cell of point v is within spacing*sqrt(3)/2 of v; with shell width
        h = spacing*sqrt(3)/(2*nshell) and k = floor(|x - v| / h), only list entries with
        |r - v| < plateaudis + (k + 1) h can lie within plateaudis of it.  The lists are
        reordered by that bound (receptor order kept among equals) and nbshell[v] holds the
        list start and the number of leading entries a bead of each shell reads, so pairs beyond the plateau
        are mostly not visited; the pairs that contribute are the same, the summation order
        differs.  Counts, starts and list memory are unchanged (host_lists and
        to_reference_grid return the reordered lists, still valid ATTRACT lists)."""
        nshell = int(nshell)
        if nshell == 0:
            self.nshell, self.nbshell, self._shell_rec = 0, None, None
            return self
        lib = _lib.load()
        V = self.dim[0] * self.dim[1] * self.dim[2]
        out = self.nblist.clone()
        nbshell = torch.empty((V, 4), dtype=torch.int32, device=self.device)
        _lib.check(lib.atj_grid_shell_lists(ctypes.byref(self.struct(whole_lists=True)),
                                            ctypes.byref(receptor.struct()), nshell, _ptr(out), _ptr(nbshell),
                                            _stream(self.device)))
        self.nblist, self.nbshell, self.nshell = out, nbshell, nshell
        self._shell_rec = receptor.beads
        return self

    def shells_valid_for(self, receptor):
        """True when the shell counts were computed for these receptor coordinates."""
        r = getattr(self, "_shell_rec", None)
        if r is None or self.nshell == 0:
            return False
        b = receptor.beads
        return r is b or (r.shape == b.shape and bool(torch.equal(r, b)))

    @classmethod
    def from_grid(cls, grid, layout=_lib.ATJ_LAYOUT_CORNER, device=None, keep_plain=False):
        """Upload a reference Grid namedtuple (read_grid / read_grid2 / pad_grid output).

        Neighbour indices must already be remapped to the receptor bead order the
        kernels will see (rec_mapping of test-grid-energrad.py:82)."""
        d = _dev(device)
        nbgrid, nblist = lists_from_grid(grid)
        inner = getattr(grid, "inner_potential_grid", None)
        outer = getattr(grid, "outer_potential_grid", None)
        if inner is not None and np.asarray(inner).ndim != 5:
            raise ValueError("the kernels need gradient grids: read the grid with read_gradients=True")
        ntypes = int(np.asarray(grid.alphabet).sum())
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(d) if a is not None else None
        return cls(origin=grid.origin, spacing=grid.gridspacing, dim=grid.dim, dim2=grid.dim2,
                   gridextension=grid.gridextension, plateaudis=grid.plateaudis, neighbourdis=grid.neighbourdis,
                   ntypes=ntypes, inner_plain=t(inner), outer_plain=t(outer),
                   nbgrid=torch.from_numpy(nbgrid).to(d),
                   nblist=torch.from_numpy(nblist.view(np.int16)).to(d),
                   layout=layout, device=d, keep_plain=keep_plain)

    @classmethod
    def build(cls, receptor, ff, lig_fftype, *, origin, spacing, dim, dim2=None, gridextension=0,
              plateaudis, neighbourdis, cutoff=50.0, layout=_lib.ATJ_LAYOUT_CORNER, device=None,
              keep_plain=False, potentials=True):
        """Compute potentials (+ force channel) and neighbour lists on the device.

        Grid semantics follow test-calc-grid-energy.py:90-196: inner point (i,j,k) sits
        at origin + spacing*(i,j,k); outer point at origin + spacing*(2*(i,j,k) - gridextension);
        the potential of ligand grid type t sums nonbon over receptor beads within
        `cutoff`, capped at the plateau value inside plateaudis; lists hold the beads
        strictly within neighbourdis of each inner point, in receptor order."""
        return cls.build_begin(receptor, ff, lig_fftype, origin=origin, spacing=spacing, dim=dim, dim2=dim2,
                               gridextension=gridextension, plateaudis=plateaudis, neighbourdis=neighbourdis,
                               cutoff=cutoff, layout=layout, device=device, keep_plain=keep_plain,
                               potentials=potentials).finish()

    @classmethod
    def build_begin(cls, receptor, ff, lig_fftype, *, origin, spacing, dim, dim2=None, gridextension=0,
                    plateaudis, neighbourdis, cutoff=50.0, layout=_lib.ATJ_LAYOUT_CORNER, device=None,
                    keep_plain=False, potentials=True):
        """The first half of build(), on the current stream without waiting for the device:
        potentials and list counts are launched and the list total starts its copy to the
        host.  Returns a GridBuild: ready() tells whether finish() can size the lists
        without waiting; finish() launches the lists and returns the grid."""
        lib = _lib.load()
        d = _dev(device)
        ntypes = len(lig_fftype)
        self = cls.__new__(cls)
        self.device = d
        self.origin = np.asarray(origin, np.float32)
        self.spacing = float(np.float32(spacing))
        self.dim = [int(v) for v in dim]
        self.dim2 = [int(v) for v in (dim2 if dim2 is not None else (2, 2, 2))]
        self.gridextension = float(gridextension)
        self.plateaudis = float(plateaudis)
        self.plateaudis2 = float(np.float32(float(plateaudis) ** 2))
        self.neighbourdis = float(neighbourdis)
        self.ntypes = ntypes
        self.layout = layout
        self.nshell = 0
        self.nbshell = None
        self.inner = self.outer = None
        self.inner_plain = self.outer_plain = None
        V = self.dim[0] * self.dim[1] * self.dim[2]
        self.nbgrid = torch.zeros((V, 2), dtype=torch.int32, device=d)
        self.nblist = torch.full((LINE,), -1, dtype=torch.int16, device=d)  # replaced by the lists in finish()
        geom = self.struct()
        st = _stream(d)
        rstruct = receptor.struct()
        if potentials:
            lt = _upload(np.asarray(lig_fftype, np.int32), d)
            inner = torch.empty((ntypes, *self.dim, 4), dtype=torch.float32, device=d)
            outer = torch.empty((ntypes, *self.dim2, 4), dtype=torch.float32, device=d) if dim2 is not None else None
            _lib.check(lib.atj_grid_build_potentials(ctypes.byref(geom), ctypes.byref(ff.struct()),
                                                     ctypes.byref(rstruct), _ptr(lt), float(cutoff),
                                                     _ptr(inner), _ptr(outer), st))
            self.inner_plain, self.outer_plain = inner, outer
            self.inner = self._relayout(inner, self.dim)
            self.outer = self._relayout(outer, self.dim2) if outer is not None else None
            if not keep_plain and layout != _lib.ATJ_LAYOUT_PLAIN:
                self.inner_plain = self.outer_plain = None
        counts = torch.empty(V, dtype=torch.int32, device=d)
        _lib.check(lib.atj_grid_count_neighbours(ctypes.byref(geom), ctypes.byref(rstruct),
                                                 float(neighbourdis), _ptr(counts), st))
        c64 = counts.to(torch.int64)
        slot = torch.where(c64 > 0, (c64 + LINE - 1) // LINE * LINE, torch.zeros_like(c64))
        ends = torch.cumsum(slot, 0) + LINE
        starts = (ends - slot).contiguous()
        total = torch.full((1,), LINE, dtype=torch.int64).pin_memory()
        if V:
            total.copy_(ends[-1:], non_blocking=True)
        counted = torch.cuda.Event()
        counted.record(torch.cuda.current_stream(d))
        return GridBuild(self, receptor, float(neighbourdis), counts, starts, total, counted,
                         torch.cuda.current_stream(d))

    def _finish_lists(self, receptor, neighbourdis, counts, starts, total):
        lib = _lib.load()
        d = self.device
        V = self.dim[0] * self.dim[1] * self.dim[2]
        self.nblist = torch.empty(total, dtype=torch.int16, device=d)
        self.nblist[:LINE].fill_(-1)
        self.nbgrid = torch.empty((V, 2), dtype=torch.int32, device=d)
        _lib.check(lib.atj_grid_fill_neighbours(ctypes.byref(self.struct()), ctypes.byref(receptor.struct()),
                                                float(neighbourdis), _ptr(counts), _ptr(starts),
                                                _ptr(self.nbgrid), _ptr(self.nblist), _stream(d)))
        return self

    def host_lists(self):
        """(counts (dx,dy,dz), 0-based starts, flat uint16 list) copied back to the host."""
        ng = self.nbgrid.cpu().numpy()
        return ng[:, 0].reshape(self.dim), ng[:, 1], self.nblist.cpu().numpy().view(np.uint16)

    def to_reference_grid(self, alphabet=None, natoms=0):
        """Host Grid namedtuple in read_grid's CSR form (needs keep_plain=True)."""
        from collections import namedtuple
        counts, starts, flat = self.host_lists()
        c = counts.reshape(-1).astype(np.int64)
        within = np.arange(c.sum()) - np.repeat(np.cumsum(c) - c, c)
        neighbours = flat[np.repeat(starts.astype(np.int64), c) + within].astype(np.uint32)
        csr_start = np.concatenate(([0], np.cumsum(c)[:-1]))
        ng = np.stack([c, np.where(c > 0, csr_start + 1, 0)], axis=1).astype(np.int32).reshape(*self.dim, 2)
        if alphabet is None:
            alphabet = np.zeros(99, bool)
            alphabet[:self.ntypes] = True
        d = dict(gridspacing=float(self.spacing), gridextension=int(self.gridextension),
                 plateaudis=self.plateaudis, neighbourdis=self.neighbourdis, alphabet=alphabet,
                 origin=self.origin.astype(np.float64), dim=np.array(self.dim), dim2=np.array(self.dim2),
                 natoms=natoms, neighbours=neighbours)
        if self.inner_plain is not None:
            d["inner_potential_grid"] = self.inner_plain.cpu().numpy()
        d["neighbour_grid"] = ng
        if self.outer_plain is not None:
            d["outer_potential_grid"] = self.outer_plain.cpu().numpy()
        return namedtuple("Grid", d.keys())(*d.values())


class GridBuild:
    """A grid build between DeviceGrid.build_begin and finish (see there)."""

    def __init__(self, grid, receptor, neighbourdis, counts, starts, total, counted, stream):
        self.grid, self.receptor, self.neighbourdis = grid, receptor, neighbourdis
        self.counts, self.starts, self.total, self.counted, self.stream = counts, starts, total, counted, stream

    def ready(self):
        """True once the list total has reached the host (finish() then waits for nothing)."""
        return self.counted.query()

    def finish(self):
        """Size and fill the neighbour lists on the build's stream; returns the DeviceGrid."""
        self.counted.synchronize()
        with torch.cuda.stream(self.stream):
            return self.grid._finish_lists(self.receptor, self.neighbourdis, self.counts, self.starts,
                                           int(self.total[0]))
