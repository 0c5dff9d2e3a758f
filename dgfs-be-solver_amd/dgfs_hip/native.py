"""Solution files in the reference's native layout (.frfss).

Follows frfs/writers/native.py:13-186 and frfs/readers/native.py:13-77: a
flat map of named entries, `<prefix>_<etype>_p<rank>` arrays of shape
(nupts, nvars, neles) plus the string entries `config` (the INI text),
`stats` (an INI with [data] fields/prefix and
[solver-time-integrator] tcurr/nsteps) and `mesh_uuid`.

The container is HDF5 when h5py is importable (the reference's format);
otherwise the same entries go into an uncompressed numpy zip archive
(one .npy member per entry, strings as 0-d byte arrays).  NativeReader
opens either, telling them apart by the file signature, and never unpickles.
"""
import os
import re
from collections import OrderedDict
from collections.abc import Mapping

import numpy as np

_HDF5_SIG = b"\x89HDF\r\n\x1a\n"
_ZIP_SIG = b"PK\x03\x04"


def _h5py():
    try:
        import h5py
    except ImportError:
        return None
    return h5py


def write_native(path, data, metadata, container="auto"):
    """Write arrays `data` {name: ndarray} and strings `metadata` to path.
    container: 'auto' (HDF5 if h5py is importable), 'hdf5' or 'npz'."""
    h5 = _h5py()
    if container == "hdf5" and h5 is None:
        raise RuntimeError("HDF5 solution files need h5py, which is not importable")
    if container not in ("auto", "hdf5", "npz"):
        raise ValueError(f"unknown container {container!r}")
    meta = {k: np.array(v, dtype="S") for k, v in metadata.items()}
    clash = set(meta) & set(data)
    if clash:
        raise ValueError(f"entries named twice: {sorted(clash)}")
    tmp = path + ".tmp"
    if h5 is not None and container != "npz":
        with h5.File(tmp, "w") as f:
            for k, v in {**data, **meta}.items():
                f[k] = v
    else:
        with open(tmp, "wb") as fh:
            np.savez(fh, **{k: np.ascontiguousarray(v) for k, v in data.items()}, **meta)
    os.replace(tmp, path)   # a reader never sees a half-written file
    return path


class NativeReader(Mapping):
    """Read-only map of a solution file (readers/native.py:13-77); scalar
    byte strings come back as str."""

    def __init__(self, fname):
        self.fname = os.path.abspath(fname)
        with open(fname, "rb") as fh:
            sig = fh.read(8)
        if sig == _HDF5_SIG:
            h5 = _h5py()
            if h5 is None:
                raise RuntimeError(f"{fname} is HDF5 and h5py is not importable")
            self._file = h5.File(fname, "r")
            self._kind = "hdf5"
        elif sig[:4] == _ZIP_SIG:
            self._file = np.load(fname, allow_pickle=False)
            self._kind = "npz"
        else:
            raise ValueError(f"{fname} is not a solution file")

    def close(self):
        self._file.close()

    def __contains__(self, aname):
        return aname in self._file

    def __getitem__(self, aname):
        if aname not in self._file:
            raise KeyError(aname)
        ret = self._file[aname]
        if self._kind == "hdf5":
            ret = ret[()] if ret.shape == () else np.array(ret)
        elif ret.shape == ():
            ret = ret[()]
        return ret.decode() if isinstance(ret, bytes) else ret

    def __iter__(self):
        return iter(self._file.keys() if self._kind == "hdf5" else self._file.files)

    def __len__(self):
        return len(list(iter(self)))

    def array_info(self, prefix):
        """{name: (etype, shape)} of the arrays under prefix, ordered by
        partition then element type (readers/native.py:44-66)."""
        names = [n for n in self if n.startswith(prefix + "_")]
        ftypes = sorted({n.split("_")[1] for n in names})
        fmaxpn = max(int(re.search(r"\d+$", n).group(0)) for n in names)
        info = OrderedDict()
        for i in range(fmaxpn + 1):
            for et in ftypes:
                n = f"{prefix}_{et}_p{i}"
                if n in self:
                    info[n] = (et, self[n].shape)
        return info

    def partition_info(self, prefix):
        """{etype: [element count per partition]} (readers/native.py:68-77)."""
        ai = self.array_info(prefix)
        npr = max(int(re.search(r"\d+$", k).group(0)) for k in ai) + 1
        nep = {v[0]: [0] * npr for v in ai.values()}
        for k, (et, shape) in ai.items():
            nep[et][int(re.search(r"\d+$", k).group(0))] = shape[-1]
        return nep


class NativeWriter:
    """Numbered solution files (writers/native.py:13-110): basename may hold
    {n} (output counter, restored from the files already in basedir on a
    restart) and {t} (the time)."""

    def __init__(self, basedir, basename, prefix, extn=".frfss", isrestart=False, container="auto"):
        self.basedir = basedir
        self.basename = basename if basename.endswith(extn) else basename + extn
        self.prefix = prefix
        self.container = container
        self.nout = self._restore_nout() if isrestart else 0

    def _restore_nout(self):
        nout = 0
        if re.search(r"{n[^}]*}", self.basename):
            bn = re.escape(self.basename)
            bn = re.sub(r"\\{n[^}]*\\}", r"(\\s*\\d+\\s*)", bn)
            bn = re.sub(r"\\{t[^}]*\\}", r"(?:.*?)", bn) + "$"
            for f in os.listdir(self.basedir):
                m = re.match(bn, f)
                if m:
                    nout = max(nout, int(m.group(1)) + 1)
        return nout

    def name_for(self, etype, prank=0):
        return f"{self.prefix}_{etype}_p{prank}"

    def write(self, arrays, metadata, tcurr):
        """arrays: {etype: ndarray (nupts, nvars, neles)} of rank 0."""
        path = os.path.join(self.basedir, self.basename.format(t=tcurr, n=self.nout))
        data = {self.name_for(et): a for et, a in arrays.items()}
        write_native(path, data, metadata, self.container)
        self.nout += 1
        return path


def dist_fields(vsize, nspcs=None):
    """Field names of a distribution file: f_0..f_{vsize-1}, and for several
    species the same names with ':<species number>' (dgfsdistwriterbi.py:31-36)."""
    base = [f"f_{i}" for i in range(vsize)]
    if nspcs is None:
        return ",".join(base)
    return ",".join(f"{f}:{p + 1}" for p in range(nspcs) for f in base)


def soln_for_system(initsoln, cfg_now, ele, fields, etype):
    """Solution array of a restart interpolated onto the current element
    basis (solvers/dgfs/system.py:79-108): the stored fields must equal
    `fields`, the stored points are those of the solution's own order and
    point rules."""
    from dgfs_hip.inifile import Inifile
    from dgfs_hip.shapes import get_shape
    solncfg = Inifile(initsoln["config"])
    solnsts = Inifile(initsoln["stats"])
    solnfields = solnsts.get("data", "fields", "")
    if not solnfields or solnfields != fields:
        raise RuntimeError("Invalid solution for system")
    soln = np.asarray(initsoln[f"soln_{etype}_p0"], dtype=float)
    upts_rule = solncfg.get(f"solver-elements-{etype}", "soln-pts",
                            "gauss-legendre" if etype == "quad" else "williams-shunn")
    fpts_rule = solncfg.get("solver-interfaces-line", "flux-pts", "gauss-legendre")
    solnb = get_shape(etype, solncfg.getint("solver", "order"), upts_rule, fpts_rule)
    if soln.shape[0] != solnb.nupts or soln.shape[-1] != ele.neles:
        raise RuntimeError("Invalid solution for mesh")
    same = solnb.nupts == ele.nupts and np.array_equal(solnb.upts, ele.basis.upts)
    if same:
        return soln
    interp = solnb.ubasis.at(ele.basis.upts)       # (nupts_now, nupts_soln)
    out = interp @ soln.reshape(solnb.nupts, -1)
    return out.reshape(ele.nupts, soln.shape[1], soln.shape[2])


def restart_stats(initsoln):
    """(tcurr, nsteps) stored by the writer (integrators/base.py:35-38)."""
    from dgfs_hip.inifile import Inifile
    st = Inifile(initsoln["stats"])
    return (st.getfloat("solver-time-integrator", "tcurr"),
            st.getint("solver-time-integrator", "nsteps", 0))
