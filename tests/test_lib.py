"""The C-ABI library builds, loads without a GPU and exports every entry point
declared in include/getm_amd.h (no compute call is made here)."""
import ctypes
import os
import re

from getm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_in_headers():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?(?:void|int|double|float)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build first: python -m getm_amd.build"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_in_headers()
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(names) == sorted(_lib.declared_symbols())


def test_create_rejects_bad_scheme_without_gpu():
    lib = _lib.load()
    # object creation that needs no device memory must still validate its input
    assert lib.create_grid(1, 0, 1, 4, 1, 3, 2, 2) is None
    g = lib.create_grid(1, 5, 1, 4, 1, 3, 2, 2)
    assert g
    assert lib.grid_set_array(g, b"nonsense", None) == 1
    assert lib.advection_create(0, g) is None
    assert lib.advection_create(7, g) is None
    lib.grid_finalize(g)


def test_create_grid_size_limits_without_gpu():
    """A layer's FP64 field must stay below 4 GiB (32-bit in-layer byte offsets)
    and each padded extent below 2^24 (24-bit index multiplies); sizes up to the
    limit are accepted."""
    lib = _lib.load()
    # 23170^2 * 8 B < 2^32 <= 23171^2 * 8 B (padded extents, halos of 2 included)
    g = lib.create_grid(1, 23166, 1, 23166, 1, 40, 2, 2)
    assert g
    lib.grid_finalize(g)
    assert lib.create_grid(1, 23167, 1, 23166, 1, 40, 2, 2) is None
    # long thin domains: padded extent 2^24 - 1 is accepted, 2^24 is not
    g = lib.create_grid(1, (1 << 24) - 5, 1, 1, 1, 1, 2, 2)
    assert g
    lib.grid_finalize(g)
    assert lib.create_grid(1, (1 << 24) - 4, 1, 1, 1, 1, 2, 2) is None
    assert lib.create_grid(1, 1, 1, (1 << 24) - 4, 1, 1, 2, 2) is None
    assert lib.create_grid(1, 4, 1, 4, 1, 3, -1, 2) is None


def test_w_momentum_rejects_bad_range_without_gpu():
    lib = _lib.load()
    # imin = 1 (Fortran) would read column -1; checked before anything is launched
    assert lib.c_w_momentum_3d(10, 8, 3, 1, 9, 2, 7, None, None, None, None, None, None, None, None, 1.0, None,
                               None) == 1


def test_vertical_advection_to_sources_rejects_bad_arguments_without_gpu():
    lib = _lib.load()
    # checked before any device call
    assert lib.c_vertical_advection_to_sources(0, 8, 3, 2, 2, None, None, None, None, None, None) == 1
    assert lib.c_vertical_advection_to_sources(10, 8, 3, -1, 2, None, None, None, None, None, None) == 1
    assert lib.c_vertical_advection_to_sources(10, 8, 3, 2, 2, None, None, None, None, None, None) == 1


def test_depth_helpers_reject_bad_arguments_without_gpu():
    lib = _lib.load()
    # checked before any device call; n = 0 launches nothing
    assert lib.c_elevation2depth(-1, None, None, 0.0, None, None, None) == 1
    assert lib.c_elevation2depth(5, None, None, 0.0, None, None, None) == 1
    assert lib.c_elevation2depth(0, None, None, 0.0, None, None, None) == 0
    assert lib.c_thickness2vertical_coordinates(0, 4, 3, None, None, None, None, None, None) == 1
    assert lib.c_thickness2vertical_coordinates(4, 4, 3, None, None, None, None, None, None) == 1


def test_interpolation_and_transport_change_reject_bad_arguments_without_gpu():
    import ctypes

    lib = _lib.load()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)  # never dereferenced: checked first
    assert lib.grid_interp_x(0, 4, 3, a, b, 0, None) == 1
    assert lib.grid_interp_x(4, 4, 3, a, a, 0, None) == 1  # in place would differ from the sequential loop
    assert lib.grid_interp_x(4, 4, 3, a, b, 2, None) == 1
    assert lib.grid_interp_y(4, 4, 3, None, b, 0, None) == 1
    assert lib.grid_interp_z(4, 4, 6, 4, a, b, 0, None) == 1  # level 5 of 4 would be written
    assert lib.grid_interp_xy(8, 6, 7, 6, 2, a, b, 1, 0, None) == 1  # column 8 of 7
    assert lib.grid_interp_xy(8, 6, 8, 5, 2, a, b, 0, 1, None) == 1  # row 6 of 5
    assert lib.grid_interp_x(1, 4, 3, a, b, 0, None) == 0  # nothing to average, nothing launched
    assert lib.c_reconstruct_transport_change(5, 4, 2, 1, 5, 0, 4, a, a, a, 1.0, None) == 1
    assert lib.c_reconstruct_transport_change(5, 4, 2, 1, 5, 1, 5, a, a, a, 1.0, None) == 1
    assert lib.c_reconstruct_transport_change(5, 4, 2, 1, 5, 1, 4, None, a, a, 1.0, None) == 1


def test_depth_update_entries_reject_bad_arguments_without_gpu():
    import ctypes

    lib = _lib.load()
    a = ctypes.c_void_p(0x1000)  # never dereferenced: checked first
    assert lib.c_alpha(-1, a, 0.1, 0.5, a, a, None) == 1
    assert lib.c_alpha(0, None, 0.1, 0.5, None, None, None) == 0
    assert lib.c_update_gvc(5, 4, 3, 1 / 3, a, 10.0, 0, a, a, a, None) == 1  # kk is 1-based
    assert lib.c_update_gvc(5, 4, 3, 1 / 3, a, 10.0, 4, a, a, a, None) == 1
    assert lib.c_update_gvc(5, 4, 3, 1 / 3, None, 10.0, 3, a, a, a, None) == 1
    assert lib.c_update_sigma(5, 0, 3, a, a, a, a, None) == 1


def test_scripts_and_modules_compile():
    """bench.py, the entry points and every package module are valid Python
    (GPU runs start them without importing them here first)."""
    import glob
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = [os.path.join(root, "bench.py"), os.path.join(root, "__graft_entry__.py")]
    files += sorted(glob.glob(os.path.join(root, "getm_amd", "*.py")))
    files += sorted(glob.glob(os.path.join(root, "oracle", "*.py")))
    for f in files:
        with open(f) as fh:
            compile(fh.read(), f, "exec")


def test_transport_set_sources_rejects_without_gpu():
    lib = _lib.load()
    assert lib.transport_set_sources(None, None) == 1


def test_velocities_from_transports_rejects_without_gpu():
    lib = _lib.load()
    assert lib.c_velocities_from_transports(-1, None, None, None, None) == 1
    assert lib.c_velocities_from_transports(8, None, None, None, None) == 1
    assert lib.c_velocities_from_transports(0, None, None, None, None) == 0


def test_header_and_ctypes_signatures_agree():
    """Every entry point declared in include/getm_amd.h is bound in getm_amd._lib
    with as many arguments as the declaration has parameters."""
    import os
    import re

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "include", "getm_amd.h")) as fh:
        text = re.sub(r"/\*.*?\*/", "", fh.read(), flags=re.S)
    decls = {}
    for m in re.finditer(r"(?:^|\n)\s*(?:const\s+)?(?:int|void|double|long)\s*\**\s*([a-zA-Z_]\w*)\s*\(([^)]*)\)\s*;",
                         text):
        params = m.group(2).strip()
        decls[m.group(1)] = 0 if params in ("", "void") else len(params.split(","))
    sigs = _lib._SIGNATURES
    assert set(decls) == set(sigs), (sorted(set(decls) ^ set(sigs)))
    for name, (_, args) in sigs.items():
        assert decls[name] == len(args), (name, decls[name], len(args))


def test_serving_drift_of_tracers_without_mass_is_finite():
    """Host logic of the serving loop: relative inventory drift, with the absolute
    change for a tracer whose admission inventory is zero (no NaN from 0/0)."""
    import torch

    from getm_amd.serve import relative_drift

    target = torch.tensor([1000.0, 0.0, -4.0], dtype=torch.float64)
    assert relative_drift(target.clone(), target) == 0.0
    assert relative_drift(torch.tensor([1000.5, 0.0, -4.0], dtype=torch.float64), target) == 5e-4
    assert relative_drift(torch.tensor([1000.0, 1e-9, -4.0], dtype=torch.float64), target) == 1e-9
    assert relative_drift(torch.tensor([1000.0, 0.0, -4.4], dtype=torch.float64), target) == abs(-4.4 + 4.0) / 4.0


def test_host_encoding_equals_device_encoding_and_decodes_within_residual():
    """encode_host (numpy, for host data) gives the bits of encode (torch), and
    decoding recovers FP64 values to 2^-16 of an FP32 ulp."""
    import torch

    import numpy as np

    from getm_amd.tracer import decode, encode, encode_host

    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.uniform(-6, 6, 5000),
                        [0.0, -0.0, 1e-39, -3e-42, 1.0, 35.0, 2.0 ** 100, np.nextafter(1.0, 2.0)]])
    hi_t, lo_t = torch.empty(x.size, dtype=torch.float32), torch.empty(x.size, dtype=torch.int16)
    encode(torch.from_numpy(x), hi_t, lo_t)
    hi, lo = encode_host(x)
    assert np.array_equal(hi.view(np.int32), hi_t.numpy().view(np.int32))
    assert np.array_equal(lo, lo_t.numpy())
    back = decode(torch.from_numpy(hi), torch.from_numpy(lo)).numpy()
    ulp = np.spacing(np.abs(hi).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(back - x) <= ulp * 2.0 ** -16 + 1e-300)
    # FP32 input is exact: zero residuals
    h32, l32 = encode_host(x.astype(np.float32))
    assert np.array_equal(h32, x.astype(np.float32)) and not l32.any()
