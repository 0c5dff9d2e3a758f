"""Matrix types with the reference's AoSoA packing (frfs/backends/base/types.py:9-208).

A 3D matrix of ioshape (nrow, nvar, narr) is stored as nrow rows of
nparr/32 blocks, each block holding nvar x 32 contiguous values
(SOA_IX(a, v, nv) = ((a/32)*nv + v)*32 + a%32); rows are padded to a
leading dimension that is a multiple of 128 bytes.  A 4D ioshape
(n0, n1, nvar, narr) has n0*n1 rows.  Device storage is a torch tensor of
shape (nrow, leaddim) on the HIP device -- memory only, every kernel that
touches it is a HIP kernel of libdgfs_hip.
"""
import numpy as np


class MatrixBase:
    _base_tags = set()

    def __init__(self, backend, dtype, ioshape, initval, extent, aliases, tags):
        self.backend = backend
        self.tags = self._base_tags | set(tags)
        self.dtype = np.dtype(dtype)
        self.itemsize = self.dtype.itemsize

        ldmod = backend.alignb // self.itemsize if "align" in self.tags else 1
        shape, ndim = list(ioshape), len(ioshape)
        if ndim == 2:
            nrow, ncol = shape
            aosoashape = shape
        else:
            nvar, narr, k = shape[-2], shape[-1], backend.soasz
            nparr = narr - narr % -k
            nrow = shape[0] if ndim == 3 else shape[0] * shape[1]
            ncol = nvar * nparr
            aosoashape = shape[:-2] + [nparr // k, nvar, k]

        self.nrow, self.ncol = int(nrow), int(ncol)
        self.datashape = aosoashape
        self.ioshape = tuple(ioshape)
        self.leaddim = self.ncol - (self.ncol % -ldmod)
        self.pitch = self.leaddim * self.itemsize
        self.nbytes = self.nrow * self.pitch
        self.traits = (self.nrow, self.leaddim, self.dtype)

        if aliases is not None:
            self.data = aliases.data.view(-1)[: self.nrow * self.leaddim].view(self.nrow, self.leaddim)
        else:
            self.data = backend._alloc((self.nrow, self.leaddim), self.dtype)
        if initval is not None:
            initval = np.asarray(initval)
            if initval.shape != self.ioshape:
                raise ValueError("Invalid initial value")
            self.set(initval)

    @property
    def _as_parameter_(self):
        return self.data.data_ptr()

    @property
    def ptr(self):
        return self.data.data_ptr()

    def _pack(self, ary):
        if ary.ndim > 2:
            n, k = ary.shape[-1], self.backend.soasz
            ary = np.pad(ary, [(0, 0)] * (ary.ndim - 1) + [(0, -n % k)], mode="constant")
            ary = ary.reshape(ary.shape[:-1] + (-1, k)).swapaxes(-2, -3)
            ary = ary.reshape(self.nrow, self.ncol)
        return np.ascontiguousarray(ary, dtype=self.dtype)

    def _unpack(self, ary):
        if len(self.ioshape) > 2:
            ary = ary.reshape(self.datashape)
            ary = ary.swapaxes(-2, -3)
            ary = ary.reshape(self.ioshape[:-1] + (-1,))
            ary = ary[..., : self.ioshape[-1]]
        return ary

    def get(self):
        # the solver may run on its own stream: wait for all of the device's work
        self._sync()
        host = self.data.detach().cpu().numpy()[:, : self.ncol]
        return self._unpack(np.ascontiguousarray(host))

    def set(self, ary):
        import torch
        packed = self._pack(np.asarray(ary))
        buf = np.zeros((self.nrow, self.leaddim), dtype=self.dtype)
        buf[:, : self.ncol] = packed
        self.data.copy_(torch.from_numpy(buf))
        self._sync()

    def _sync(self):
        sync = getattr(self.backend, "synchronize", None)
        if sync is not None:
            sync()
        else:
            import torch
            torch.cuda.synchronize(self.data.device)


class Matrix(MatrixBase):
    _base_tags = {"dense"}

    def __init__(self, backend, ioshape, initval, extent, aliases, tags):
        super().__init__(backend, backend.fpdtype, ioshape, initval, extent, aliases, tags)


class ConstMatrix(MatrixBase):
    _base_tags = {"const", "dense"}

    def __init__(self, backend, initval, extent, tags):
        initval = np.asarray(initval)
        super().__init__(backend, backend.fpdtype, initval.shape, initval, extent, None, tags)


class MatrixBank(list):
    """A list of matrices with one marked active (types.py:170-208)."""

    def __init__(self, backend, mats, initbank=0):
        super().__init__(mats)
        self.backend = backend
        self._curr_idx = initbank

    def __getattr__(self, attr):
        return getattr(self[self._curr_idx], attr)

    @property
    def active(self):
        return self._curr_idx

    @active.setter
    def active(self, idx):
        self._curr_idx = idx


class View:
    """One side of a set of interface points (the reference's view of the
    flux-point matrix, backends/base/types.py:211-264): interface point i of
    variable v lives at base + offsets[i] + 32 v."""

    def __init__(self, backend, base, offsets):
        self.backend = backend
        self.base = base
        offsets = np.asarray(offsets, dtype=np.int64).reshape(-1)
        if offsets.size and (offsets.min() < 0 or offsets.max() >= base.nrow * base.leaddim):
            raise ValueError("view offsets outside the base matrix")
        self.n = int(offsets.size)
        self.offsets = backend.const_array(offsets if offsets.size else np.zeros(1, np.int64), np.int64)
