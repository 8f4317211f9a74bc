"""HDF5 field I/O (libstencilfft_hdf5.so) on the CPU.

The file format is the reference's (zmodel/utils.cu:88-281): an N^d dataset of
the compound {"real", "imag"} doubles in global natural order, each rank's
block-cyclic hyperslab vector-packed in memory.  The files are checked with
an independent reader and writer that call libhdf5 directly (ctypes, whole
dataset, no hyperslabs), against the oracle's natural <-> packed maps
(zmodel/utils.h:99-320).
"""
import ctypes
import os

import numpy as np
import pytest

SHAPES = [(2, 16, 2, 2), (2, 32, 2, 4), (2, 64, 4, 2), (2, 64, 2, 8), (3, 16, 2, 2), (3, 32, 2, 4)]


def _hdf5_prefix():
    from stencilfft_amd import build
    return build.hdf5_prefix()


pytestmark = pytest.mark.skipif(_hdf5_prefix() is None, reason="no HDF5 C library installed")


@pytest.fixture(scope="module")
def h5io():
    from stencilfft_amd import build, hdf5
    if not os.path.exists(hdf5.HDF5_LIB_PATH):
        build.build_hdf5()
    return hdf5


class H5:
    """The few libhdf5 calls an independent whole-dataset reader/writer needs."""

    hid = ctypes.c_int64
    hsize = ctypes.c_ulonglong

    def __init__(self):
        lib = ctypes.CDLL(os.path.join(_hdf5_prefix(), "lib", "libhdf5.so"))
        assert lib.H5open() >= 0
        self.lib = lib
        hid = self.hid
        for name, res, args in [
            ("H5Fcreate", hid, [ctypes.c_char_p, ctypes.c_uint, hid, hid]),
            ("H5Fopen", hid, [ctypes.c_char_p, ctypes.c_uint, hid]),
            ("H5Fclose", ctypes.c_int, [hid]),
            ("H5Dcreate2", hid, [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]),
            ("H5Dopen2", hid, [hid, ctypes.c_char_p, hid]),
            ("H5Dclose", ctypes.c_int, [hid]),
            ("H5Dget_space", hid, [hid]),
            ("H5Dget_type", hid, [hid]),
            ("H5Dread", ctypes.c_int, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
            ("H5Dwrite", ctypes.c_int, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
            ("H5Screate_simple", hid, [ctypes.c_int, ctypes.POINTER(self.hsize), ctypes.c_void_p]),
            ("H5Sget_simple_extent_ndims", ctypes.c_int, [hid]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [hid, ctypes.POINTER(self.hsize), ctypes.c_void_p]),
            ("H5Sclose", ctypes.c_int, [hid]),
            ("H5Tcreate", hid, [ctypes.c_int, ctypes.c_size_t]),
            ("H5Tinsert", ctypes.c_int, [hid, ctypes.c_char_p, ctypes.c_size_t, hid]),
            ("H5Tget_class", ctypes.c_int, [hid]),
            ("H5Tget_size", ctypes.c_size_t, [hid]),
            ("H5Tget_nmembers", ctypes.c_int, [hid]),
            ("H5Tget_member_name", ctypes.c_void_p, [hid, ctypes.c_uint]),
            ("H5Tget_member_offset", ctypes.c_size_t, [hid, ctypes.c_uint]),
            ("H5free_memory", ctypes.c_int, [ctypes.c_void_p]),
            ("H5Tclose", ctypes.c_int, [hid]),
        ]:
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        self.double = hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value

    def ctype(self):
        t = self.lib.H5Tcreate(6, 16)  # H5T_COMPOUND
        assert self.lib.H5Tinsert(t, b"real", 0, self.double) >= 0
        assert self.lib.H5Tinsert(t, b"imag", 8, self.double) >= 0
        return t

    def read(self, filename, dataset):
        """(natural-order complex array, [(member name, offset)], class, size)."""
        L = self.lib
        f = L.H5Fopen(os.fsencode(filename), 0, 0)
        assert f >= 0
        d = L.H5Dopen2(f, dataset.encode(), 0)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (self.hsize * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        ft = L.H5Dget_type(d)
        members = []
        for i in range(L.H5Tget_nmembers(ft)):
            p = L.H5Tget_member_name(ft, i)
            members.append((ctypes.string_at(p).decode(), L.H5Tget_member_offset(ft, i)))
            L.H5free_memory(p)
        info = (L.H5Tget_class(ft), L.H5Tget_size(ft))
        out = np.empty(tuple(dims), dtype=np.complex128)
        mt = self.ctype()
        assert L.H5Dread(d, mt, 0, 0, 0, out.ctypes.data) >= 0  # H5S_ALL, H5S_ALL, H5P_DEFAULT
        for h, close in [(mt, L.H5Tclose), (ft, L.H5Tclose), (sp, L.H5Sclose), (d, L.H5Dclose), (f, L.H5Fclose)]:
            close(h)
        return out, members, info

    def write(self, filename, dataset, nat):
        L = self.lib
        nat = np.ascontiguousarray(nat, dtype=np.complex128)
        f = L.H5Fcreate(os.fsencode(filename), 2, 0, 0)  # H5F_ACC_TRUNC
        assert f >= 0
        dims = (self.hsize * nat.ndim)(*nat.shape)
        sp = L.H5Screate_simple(nat.ndim, dims, None)
        t = self.ctype()
        d = L.H5Dcreate2(f, dataset.encode(), t, sp, 0, 0, 0)
        assert d >= 0
        assert L.H5Dwrite(d, t, 0, 0, 0, nat.ctypes.data) >= 0
        for h, close in [(d, L.H5Dclose), (t, L.H5Tclose), (sp, L.H5Sclose), (f, L.H5Fclose)]:
            close(h)


@pytest.fixture(scope="module")
def h5():
    return H5()


def _random_natural(dims, n, seed):
    rng = np.random.default_rng(seed)
    shape = (n,) * dims
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_library_exports_every_declared_symbol(h5io):
    import re
    from conftest import ROOT
    text = open(os.path.join(ROOT, "include", "stencilfft_hdf5.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    syms = sorted(set(re.findall(r"\b(zm_[a-z0-9_]+)\s*\(", text)))
    assert syms == sorted(h5io.SIGNATURES)
    lib = h5io._lib()
    assert all(hasattr(lib, s) for s in syms)


@pytest.mark.parametrize("dims,n,p,b", SHAPES)
def test_store_writes_natural_order(h5io, h5, oracle_mod, tmp_path, dims, n, p, b):
    """store_host_hdf5 of the packed grid == the natural-order field, read whole."""
    nat = _random_natural(dims, n, seed=n + p + b + dims)
    packed = oracle_mod.Shape(dims, n, p, b).natural_to_real(nat)
    fn = str(tmp_path / "field.h5")
    h5io.store_host_hdf5((dims, n, p, b), packed, fn, "u")
    got, members, (cls, size) = h5.read(fn, "u")
    assert got.shape == (n,) * dims
    assert members == [("real", 0), ("imag", 8)] and cls == 6 and size == 16
    np.testing.assert_array_equal(got, nat)


@pytest.mark.parametrize("dims,n,p,b", SHAPES)
def test_load_reads_rank_hyperslabs(h5io, h5, oracle_mod, tmp_path, dims, n, p, b):
    """init_host_hdf5 of a natural-order file == the oracle's packed grid."""
    nat = _random_natural(dims, n, seed=7 * n + b)
    fn = str(tmp_path / "field.h5")
    h5.write(fn, "field", nat)
    got = h5io.init_host_hdf5((dims, n, p, b), fn, "field")
    np.testing.assert_array_equal(got, oracle_mod.Shape(dims, n, p, b).natural_to_real(nat))


@pytest.mark.parametrize("dims,n,p,b", [(2, 64, 4, 4), (3, 32, 2, 2)])
def test_round_trip_and_init_host(h5io, oracle_mod, tmp_path, dims, n, p, b):
    """store then load is the identity; the reference's default field survives it."""
    grid = oracle_mod.Shape(dims, n, p, b).init_host()
    fn = str(tmp_path / "init.h5")
    h5io.store_host_hdf5((dims, n, p, b), grid, fn, "data")
    np.testing.assert_array_equal(h5io.init_host_hdf5((dims, n, p, b), fn, "data"), grid)
    # a file written for one (P, B) reads into another distribution of the same N
    other = (dims, n, 2, 1)
    np.testing.assert_array_equal(h5io.init_host_hdf5(other, fn, "data"),
                                  oracle_mod.Shape(*other).natural_to_real(oracle_mod.Shape(dims, n, p, b)
                                                                           .real_to_natural(grid)))


def test_errors(h5io, h5, tmp_path):
    from stencilfft_amd import ZmError
    from stencilfft_amd._lib import ZM_ERR_IO, ZM_ERR_SHAPE
    fn = str(tmp_path / "f.h5")
    with pytest.raises(ZmError) as e:
        h5io.init_host_hdf5((2, 16, 2, 2), str(tmp_path / "missing.h5"), "data")
    assert e.value.status == ZM_ERR_IO
    h5.write(fn, "data", _random_natural(2, 16, 0))
    with pytest.raises(ZmError) as e:
        h5io.init_host_hdf5((2, 16, 2, 2), fn, "nope")
    assert e.value.status == ZM_ERR_IO
    with pytest.raises(ZmError) as e:  # a 16x16 dataset read as a 32x32 field
        h5io.init_host_hdf5((2, 32, 2, 2), fn, "data")
    assert e.value.status == ZM_ERR_SHAPE
    with pytest.raises(ZmError) as e:  # a 2D dataset read as 3D
        h5io.init_host_hdf5((3, 16, 2, 2), fn, "data")
    assert e.value.status == ZM_ERR_SHAPE
    with pytest.raises(ZmError) as e:  # invalid distribution
        h5io.store_host_hdf5((2, 16, 3, 2), np.zeros(256, np.complex128), fn, "data")
    assert e.value.status == ZM_ERR_SHAPE
    with pytest.raises(ZmError) as e:  # directory that does not exist
        h5io.store_host_hdf5((2, 16, 2, 2), np.zeros(256, np.complex128), str(tmp_path / "no" / "f.h5"), "d")
    assert e.value.status == ZM_ERR_IO
    with pytest.raises(ValueError):
        h5io.store_host_hdf5((2, 16, 2, 2), np.zeros(100, np.complex128), fn, "data")


def test_cpp_facade_wrappers(h5io, tmp_path):
    """zmodel::store_host_hdf5 / init_host_hdf5 (include/zmodel.hpp) from C++,
    host code only: compiled with g++ against the two libraries."""
    import shutil
    import subprocess
    from conftest import ROOT
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    pkg = os.path.join(ROOT, "stencilfft_amd")
    exe = str(tmp_path / "facade_hdf5")
    cmd = [gxx, "-std=c++17", "-O1", os.path.join(ROOT, "tests", "facade_hdf5.cpp"), "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-L", pkg, "-lstencilfft_hdf5", "-lstencilfft",
           f"-Wl,-rpath,{pkg}", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    run = subprocess.run([exe, str(tmp_path / "ramp.h5")], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and "facade hdf5 ok" in run.stdout, (run.returncode, run.stdout, run.stderr)
