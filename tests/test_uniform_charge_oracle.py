"""The oracle's uniform-charge slicing (slicing.py:618-686) pinned by its
defining properties, as the reference's test_slicing.py states them for
this slicer: the particles inside the cuts are split into slices whose
counts differ by at most one and sum to the number inside; every particle
of slice i lies between the edges z_bins[i] and z_bins[i+1]; the counts
are those of the slice indices; the remainder goes to slices drawn by
random.sample, so a seeded generator reproduces them."""
import random

import numpy as np

from oracle import refpath as R
from oracle import workloads as W


def test_uniform_charge_properties():
    ob = W.linear_bunch(10007, seed=3)
    for kw in (dict(n_sigma_z=2), dict(z_cuts=(-0.3, 0.4)), dict()):
        random.seed(11)
        sd = R.slice_uniform_charge(ob, 37, **kw)
        tail, head = R.long_cuts(ob, kw.get('n_sigma_z'), kw.get('z_cuts'))
        inside = np.sum((ob.z >= tail) & (ob.z < head))
        cnt = sd.n_macroparticles_per_slice
        assert cnt.sum() == inside
        assert cnt.max() - cnt.min() <= 1
        np.testing.assert_array_equal(np.bincount(sd.slice_index_of_particle[
            sd.slice_index_of_particle >= 0], minlength=37), cnt)
        for i in range(37):
            z = ob.z[sd.slice_index_of_particle == i]
            assert np.all(z >= sd.z_bins[i]) and np.all(z <= sd.z_bins[i + 1])
        random.seed(11)
        sd2 = R.slice_uniform_charge(ob, 37, **kw)
        np.testing.assert_array_equal(sd2.slice_index_of_particle, sd.slice_index_of_particle)
