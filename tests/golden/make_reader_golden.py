"""Golden vectors for the grid reader.

1. Writes a small synthetic ATTRACT grid (reader_small.grid) with
   attract_jax_amd.grid.write_grid: 3 potential types (1, 5, 32), inner 5x4x3,
   outer 4x4x3 with some empty outer points, ragged neighbour lists with empty voxels.
2. Parses it with the reference's own readers -- read_grid.read_grid (with and without
   gradients), read_grid.pad_grid (list and integer padding), read_grid2.read_grid --
   imported from /root/reference, and stores every field in reader_golden.npz.
The reference code is only run here, to produce these fixtures."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from attract_jax_amd.grid import write_grid  # noqa: E402


def synthetic_grid_bytes():
    rng = np.random.default_rng(7)
    dim, dim2 = (5, 4, 3), (4, 4, 3)
    alphabet = np.zeros(99, bool)
    alphabet[[0, 4, 31]] = True
    inner = rng.normal(size=(3, *dim, 4)).astype(np.float32)
    outer = rng.normal(size=(3, *dim2, 4)).astype(np.float32)
    counts = rng.integers(0, 5, size=dim)
    counts[0, 0, 0] = 0
    counts[1, 2, 1] = 7
    neighbours = rng.integers(0, 40, size=int(counts.sum()))
    mask = np.zeros(dim2, bool)
    mask[1:3, 1:3, 1] = True
    return write_grid(gridspacing=0.9, gridextension=4, plateaudis=5.0, neighbourdis=7.0, alphabet=alphabet,
                      origin=(-2.0, -1.5, 0.25), dim=dim, dim2=dim2, natoms=40, inner_potential_grid=inner,
                      outer_potential_grid=outer, nb_counts=counts, neighbours=neighbours,
                      nb_types=rng.integers(1, 33, size=len(neighbours)), outer_mask=mask)


def main():
    data = synthetic_grid_bytes()
    open(os.path.join(HERE, "reader_small.grid"), "wb").write(data)
    sys.path.insert(0, "/root/reference")
    import read_grid as ref1  # noqa: E402
    import read_grid2 as ref2  # noqa: E402
    out = {}

    def store(prefix, g):
        out[prefix + "__fields"] = np.array(g._fields)
        for f in g._fields:
            v = getattr(g, f)
            if v is None:
                continue
            out[prefix + "__" + f] = np.asarray(v)

    g = ref1.read_grid(data, read_gradients=True)
    store("rg_grad", g)
    store("rg_nograd", ref1.read_grid(data, read_gradients=False))
    store("pad_list", ref1.pad_grid(g, [2, 1, 3]))
    store("pad_int", ref1.pad_grid(ref1.read_grid(data, read_gradients=True), 2))
    store("rg2_grad", ref2.read_grid(data, read_gradients=True))
    np.savez_compressed(os.path.join(HERE, "reader_golden.npz"), **out)
    print("wrote", len(out), "arrays;", len(data), "grid bytes")


if __name__ == "__main__":
    main()
