"""Bunch and slice monitors (PyHEADTAIL monitors/monitors.py:25-380) with
their histories kept on the device.

BunchMonitor.dump records one device reduction per step: the moment sums of
the six coordinates (phx_bunch_sums, 33 FP64 values used here) into a
device ring of buffer_size rows, with no host synchronisation.  When the
buffer is written (every write_buffer_every steps and at the last step) the
rows are copied to the host and turned into mean_*, sigma_*, epsn_* with
the arithmetic of Particles.mean_x() etc., so the stored values equal what
those methods return at the same step.  SliceMonitor keeps the slice
statistics of each step in a device ring likewise.

The file layout is the reference's: groups 'Bunch' (datasets of shape
(n_steps,)) and 'Slices' ((n_slices, n_steps)), metadata from
parameters_dict.  With h5py importable it is an HDF5 file filename + '.h5'
as in the reference; without it, a NumPy archive filename + '.npz' with
keys 'Bunch/<stat>', 'Slices/<stat>' and 'attrs/<key>', rewritten at each
buffer write.
"""
import numpy as np
import torch

from .. import _phx
from ..engine import stats_from_sums
from ..general.element import Printing

try:
    import h5py
except ImportError:  # not in every image: the .npz layout stands in
    h5py = None

BUNCH_STATS = ['mean_x', 'mean_xp', 'mean_y', 'mean_yp', 'mean_z', 'mean_dp', 'sigma_x',
               'sigma_y', 'sigma_z', 'sigma_dp', 'epsn_x', 'epsn_y', 'epsn_z',
               'macroparticlenumber']
SLICE_STATS = ['mean_x', 'mean_xp', 'mean_y', 'mean_yp', 'mean_z', 'mean_dp', 'sigma_x',
               'sigma_y', 'sigma_z', 'sigma_dp', 'epsn_x', 'epsn_y', 'epsn_z',
               'n_macroparticles_per_slice']
_NSUMS = 33  # sums, shifted sums, products of the six coordinates


class _Store(object):
    """The monitor file: HDF5 when h5py is available, else .npz."""

    def __init__(self, filename, groups, parameters_dict):
        self.h5 = h5py is not None
        self.path = filename + ('.h5' if self.h5 else '.npz')
        if self.h5:
            with h5py.File(self.path, 'w') as f:
                for k, v in (parameters_dict or {}).items():
                    f.attrs[k] = v
                for g, datasets in groups.items():
                    grp = f.create_group(g)
                    for name, shape in datasets:
                        grp.create_dataset(name, shape=shape, compression='gzip',
                                           compression_opts=9)
        else:
            self.data = {'%s/%s' % (g, name): np.zeros(shape)
                         for g, datasets in groups.items() for name, shape in datasets}
            for k, v in (parameters_dict or {}).items():
                self.data['attrs/' + k] = np.asarray(v)
            np.savez(self.path, **self.data)

    def write(self, columns):
        """columns: {(group, name): (slice of the step axis, values)}."""
        if self.h5:
            with h5py.File(self.path, 'a') as f:
                for (g, name), (sl, vals) in columns.items():
                    if vals.ndim == 1:
                        f[g][name][sl] = vals
                    else:
                        f[g][name][:, sl] = vals
        else:
            for (g, name), (sl, vals) in columns.items():
                a = self.data['%s/%s' % (g, name)]
                if vals.ndim == 1:
                    a[sl] = vals
                else:
                    a[:, sl] = vals
            np.savez(self.path, **self.data)


class Monitor(Printing):
    def dump(self, bunch):
        raise NotImplementedError


class _BunchRecorder(object):
    """Device ring of per-step moment sums plus the host scalars the
    statistics need; stats() gives the recorded steps' statistics."""

    def __init__(self, stats, buffer_size):
        self.stats, self.size = list(stats), buffer_size
        self.ring = None
        self.host = {}

    def record(self, bunch, pos):
        L = _phx.lib()
        n = bunch.macroparticlenumber
        if self.ring is None:
            self.ring = torch.zeros((self.size, L.phx_bunch_sums_len(6)), dtype=torch.float64,
                                    device=_phx.device())
        arrays = [bunch.x, bunch.xp, bunch.y, bunch.yp, bunch.z, bunch.dp]
        ptrs = (_phx.P * 6)(*[a.data_ptr() for a in arrays])
        ws = _phx.workspace(n)
        _phx.call(L.phx_bunch_sums, ptrs, 6, n, self.ring[pos].data_ptr(), ws.data_ptr(),
                  _phx.stream())
        self.host[pos] = (n, bunch.betagamma, bunch.p0)
        for s in self.stats:
            if s not in BUNCH_STATS:     # another attribute or method of the bunch
                v = getattr(bunch, s)
                self.host[(s, pos)] = v() if callable(v) else v

    def values(self, positions):
        rows = self.ring[positions, :_NSUMS].cpu().numpy()
        out = {s: np.empty(len(positions)) for s in self.stats}
        for r, pos in enumerate(positions):
            n, bg, p0 = self.host[pos]
            st = stats_from_sums(rows[r:r + 1], n, bg, p0)
            for s in self.stats:
                if s == 'macroparticlenumber':
                    out[s][r] = n
                elif s in st:
                    out[s][r] = st[s][0]
                else:
                    out[s][r] = self.host[(s, pos)]
        return out


class BunchMonitor(Monitor):
    """Bunch statistics per step to a file, buffered."""

    def __init__(self, filename, n_steps, parameters_dict=None, write_buffer_every=512,
                 buffer_size=4096, *args, **kwargs):
        self.stats_to_store = kwargs.pop('stats_to_store', list(BUNCH_STATS))
        self.filename = filename
        self.n_steps = n_steps
        self.i_steps = 0
        self.buffer_size = buffer_size
        self.write_buffer_every = write_buffer_every
        self._store = _Store(filename, {'Bunch': [(s, (n_steps,))
                                                  for s in sorted(self.stats_to_store)]},
                             parameters_dict)
        self._rec = _BunchRecorder(self.stats_to_store, buffer_size)

    def dump(self, bunch):
        self._rec.record(bunch, self.i_steps % self.buffer_size)
        if ((self.i_steps + 1) % self.write_buffer_every == 0 or
                (self.i_steps + 1) == self.n_steps):
            self._write_buffer_to_file()
        self.i_steps += 1

    def _steps_in_buffer(self):
        n_entries = min(self.i_steps + 1, self.buffer_size)
        lo = self.i_steps + 1 - n_entries
        return lo, self.i_steps + 1

    def _write_buffer_to_file(self):
        lo, hi = self._steps_in_buffer()
        pos = [s % self.buffer_size for s in range(lo, hi)]
        vals = self._rec.values(pos)
        try:
            self._store.write({('Bunch', s): (slice(lo, hi), vals[s])
                               for s in self.stats_to_store})
        except IOError:
            self.warns('Bunch monitor file is temporarily unavailable. \n')


class SliceMonitor(BunchMonitor):
    """Bunch and slice statistics per step to a file, buffered."""

    def __init__(self, filename, n_steps, slicer, parameters_dict=None, write_buffer_every=512,
                 buffer_size=4096, *args, **kwargs):
        self.bunch_stats_to_store = kwargs.pop('bunch_stats_to_store', list(BUNCH_STATS))
        self.slice_stats_to_store = kwargs.pop('slice_stats_to_store', list(SLICE_STATS))
        self.filename = filename
        self.n_steps = n_steps
        self.i_steps = 0
        self.slicer = slicer
        self.buffer_size = buffer_size
        self.write_buffer_every = write_buffer_every
        S = slicer.n_slices
        self._store = _Store(filename, {
            'Bunch': [(s, (n_steps,)) for s in self.bunch_stats_to_store],
            'Slices': [(s, (S, n_steps)) for s in self.slice_stats_to_store]}, parameters_dict)
        self._rec = _BunchRecorder(self.bunch_stats_to_store, buffer_size)
        self._slice_ring = None

    def dump(self, bunch):
        pos = self.i_steps % self.buffer_size
        slice_set = bunch.get_slices(self.slicer, statistics=True)
        self._rec.record(bunch, pos)
        if self._slice_ring is None:
            self._slice_ring = torch.zeros(
                (len(self.slice_stats_to_store), self.buffer_size, self.slicer.n_slices),
                dtype=torch.float64, device=_phx.device())
        for i, s in enumerate(self.slice_stats_to_store):
            self._slice_ring[i, pos].copy_(getattr(slice_set, s))
        if ((self.i_steps + 1) % self.write_buffer_every == 0 or
                (self.i_steps + 1) == self.n_steps):
            self._write_buffer_to_file()
        self.i_steps += 1

    def _write_buffer_to_file(self):
        lo, hi = self._steps_in_buffer()
        pos = [s % self.buffer_size for s in range(lo, hi)]
        vals = self._rec.values(pos)
        sl = self._slice_ring[:, pos].cpu().numpy()     # [stat][step][slice]
        cols = {('Bunch', s): (slice(lo, hi), vals[s]) for s in self.bunch_stats_to_store}
        for i, s in enumerate(self.slice_stats_to_store):
            cols[('Slices', s)] = (slice(lo, hi), sl[i].T)
        try:
            self._store.write(cols)
        except IOError:
            self.warns('Slice monitor file is temporarily unavailable. \n')
