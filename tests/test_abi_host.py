"""CPU tests of the C ABI library: it loads, exports every declared symbol, and its host-side
constructors (no GPU needed) build the same tables as the oracle, bit for bit."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import _lib
from feriphys_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "feriphys_gpu.h")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fp_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = fa.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fp_[a-z0-9_]+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        getattr(lib, s)
    # the Python mirror binds exactly the declared entry points
    assert set(_lib.SIGNATURES) == set(syms)


def test_c99_caller(tmp_path):
    """The boundary from plain C: tests/native/abi_c_check.c includes the header as C99, links the
    library and checks construction, readback, error codes and ownership (and a step when a GPU is
    present)."""
    exe = str(tmp_path / "abi_c_check")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "abi_c_check.c"), "-L", libdir, "-lferiphys_gpu",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip()
    assert out == "ok", out


def test_cpp_mirror_reads_like_the_reference_tests(tmp_path):
    """include/feriphys.hpp mirrors the reference's springy API in C++ over the C ABI;
    tests/native/mirror_check.cpp restates the reference's SpringyMesh unit tests (ctor_triangle,
    strip_ctor, get_vertices) through it, plus Cloth::new's counts and the panics as exceptions."""
    exe = str(tmp_path / "mirror_check")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "mirror_check.cpp"), "-L", libdir, "-lferiphys_gpu",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip()
    assert out == "ok", out


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: asked for more it says so and exits with status 2, before it
    touches a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""


def test_rust_binding_declares_every_entry_point():
    """The integration guide's appendix (the extern "C" block a Rust caller adds) declares each
    function of the header once, with the header's parameter count."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = doc[doc.index("## Appendix: every entry point"):]
    rust = re.findall(r"pub fn (fp_[a-z0-9_]+)\(([^)]*)\)", block)
    names = [n for n, _ in rust]
    assert sorted(names) == declared_symbols() and len(names) == len(set(names))
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, args in rust:
        c_args = re.search(r"\b" + name + r"\s*\(([^)]*)\)", txt).group(1).strip()
        n_c = 0 if c_args in ("", "void") else c_args.count(",") + 1
        n_r = 0 if not args.strip() else args.count(",") + 1
        assert n_c == n_r, name


def test_host_alloc_paths_without_a_device():
    """fp_host_alloc's checks that need no GPU: a NULL out pointer is FP_EINVAL, zero bytes give
    NULL and FP_OK, and fp_host_free(NULL) is a no-op."""
    L = _lib.lib()
    assert L.fp_host_alloc(16, None) == _lib.FP_EINVAL
    p = C.c_void_p(12345)
    assert L.fp_host_alloc(0, C.byref(p)) == _lib.FP_OK and not p.value
    assert L.fp_host_free(None) == _lib.FP_OK
    assert fa.pinned_empty((0, 10)).shape == (0, 10)


def test_config_default():
    # springy/config.rs:20-33
    c = fa.config_default()
    assert c.integration == fa.Integration.Rk4
    assert c.dt == np.float32(0.001)
    assert c.gravity == (0.0, -10.0, 0.0) and c.wind == (0.0, 0.0, 0.0)
    assert (c.lift_coefficient, c.drag_coefficient) == (1.0, 1.0)
    assert c.coefficient_of_restitution == np.float32(0.95) and c.coefficient_of_friction == np.float32(0.3)


def compare_mesh(g, o):
    assert g.counts() == o.counts()
    v01, kcl, fc = g.struts()
    so = o.struts()
    assert (v01 == so["v01"]).all()
    assert (bits(kcl) == bits(so["kcl"])).all()
    assert (fc == so["faces"]).all()
    s3, v3 = g.faces()
    os3, ov3 = o.faces()
    assert (s3 == os3).all() and (v3 == ov3).all()
    assert (bits(g.get_points()) == bits(o.points())).all()
    assert (g.pins() == o.pins()).all()
    has, kcr = g.torsion()
    assert (has == so["has_torsion"]).all()
    assert (bits(kcr[has]) == bits(so["torsion"][has])).all()


@pytest.mark.parametrize("rows,cols,plane", [(32, 32, 1), (32, 32, 0), (5, 7, 0), (2, 2, 0), (3, 9, 1), (2, 11, 1),
                                             (17, 3, 0), (64, 48, 1)])
def test_cloth_tables_match_oracle(rows, cols, plane):
    c = wl.config0(rows=rows, cols=cols).cloths[0]
    compare_mesh(fa.Cloth(*c.args(), plane=plane).mesh, orc.Mesh.cloth(*c.args(), plane=plane))


def test_cloth_1024_tables_match_oracle():
    """The benchmark cloth: 1,048,576 vertices, 5,230,601 struts (one shear diagonal per quad,
    truncated bends), 2,093,058 faces."""
    c = wl.config1().cloths[0]
    g = fa.Cloth(*c.args(), plane=c.plane).mesh
    o = orc.Mesh.cloth(*c.args(), plane=c.plane)
    assert g.counts() == (1048576, 5230601, 2093058, 1024)
    compare_mesh(g, o)


@pytest.mark.parametrize("rows,cols", [(2, 2), (3, 3), (4, 7), (10, 10), (31, 5)])
def test_cloth_counts_closed_form(rows, cols):
    """S = 3(R-1)(C-1) + (R-1) + (C-1) + 2(R-2)(C-2); F = 2(R-1)(C-1) (springy/cloth.rs:63-110)."""
    c = wl.config0(rows=rows, cols=cols).cloths[0]
    n, s, f, p = fa.Cloth(*c.args()).mesh.counts()
    assert n == rows * cols
    assert s == 3 * (rows - 1) * (cols - 1) + (rows - 1) + (cols - 1) + 2 * max(rows - 2, 0) * max(cols - 2, 0)
    assert f == 2 * (rows - 1) * (cols - 1)


@pytest.mark.parametrize("rows,cols", [(1, 5), (5, 1), (1, 1), (2, 1), (0, 4)])
def test_cloth_below_two_by_two_is_refused(rows, cols):
    """Cloth::new's bend loops run over 0..(rows - 2) and 0..(cols - 2) (springy/cloth.rs:101-102),
    which underflow and panic for a side below 2: the library returns FP_EINVAL and the oracle
    refuses the same sizes."""
    c = wl.config0(rows=2, cols=2).cloths[0]
    args = (rows, cols) + tuple(c.args())[2:-1] + ([],)
    with pytest.raises(fa.FpError) as e:
        fa.Cloth(*args)
    assert e.value.code == _lib.FP_EINVAL
    with pytest.raises(ValueError):
        orc.Mesh.cloth(*args)


def test_three_by_three_cloth_strut_list_by_hand():
    """Brute force on a tiny input: every strut of the 3x3 cloth, enumerated by hand from the
    triangle list (top-left triangles (i+1, i+3, i) of quads 0, 1, 3, 4, then bottom-right
    triangles (i-2, i+1, i) for i = 3, 4, 6, 7, which add only the last column's verticals and the
    last row's horizontals) and the two bending struts (0,2), (0,6)."""
    c = wl.config0(rows=3, cols=3).cloths[0]
    v01, kcl, _ = fa.Cloth(*c.args()).mesh.struts()
    expect = [(1, 3), (0, 3), (0, 1), (2, 4), (1, 4), (1, 2), (4, 6), (3, 6), (3, 4), (5, 7), (4, 7), (4, 5),
              (2, 5), (6, 7), (5, 8), (7, 8), (0, 2), (0, 6)]
    assert [tuple(map(int, r)) for r in v01] == expect
    s = np.float32(c.spacing)
    L = kcl[:, 2]
    assert abs(L[2] - s) < 1e-7 and abs(L[1] - s) < 1e-7         # tensile: the spacing
    assert abs(L[0] - s * np.sqrt(2)) < 1e-7                      # shear diagonal
    assert abs(L[16] - 2 * s) < 1e-7 and abs(L[17] - 2 * s) < 1e-7  # bends
    # shear struts carry the shear override, the rest the tensile/bend constants (1/L * nominal)
    assert kcl[0, 0] == np.float32(1.0) / L[0] * np.float32(c.sk)
    assert kcl[2, 0] == np.float32(1.0) / L[2] * np.float32(c.tk)
    assert kcl[13, 0] == np.float32(1.0) / L[13] * np.float32(c.tk)
    assert kcl[16, 0] == np.float32(1.0) / L[16] * np.float32(c.bk)


def test_random_mesh_tables_match_oracle():
    rng = np.random.default_rng(3)
    for trial in range(5):
        n = 40
        pos = rng.normal(size=(n, 3)).astype(np.float32)
        # a random manifold-ish strip: fan triangles with distinct corners
        tris = []
        for t in range(60):
            a, b, c = rng.choice(n, 3, replace=False)
            tris.append((a, b, c))
        try:
            o = orc.Mesh.new(pos, np.array(tris).reshape(-1), 2.5, 7.0, 0.5, torsion=(3.0, 1.0),
                             overrides=[(1, 2, 9.0, 9.5), (int(tris[0][0]), int(tris[0][1]), 4.0, 4.5)])
        except ValueError:
            with pytest.raises(ValueError):
                fa.SpringyMesh(pos, np.array(tris).reshape(-1), 2.5, 7.0, 0.5, (3.0, 1.0),
                               {(1, 2): (9.0, 9.5), (int(tris[0][0]), int(tris[0][1])): (4.0, 4.5)})
            continue
        g = fa.SpringyMesh(pos, np.array(tris).reshape(-1), 2.5, 7.0, 0.5, (3.0, 1.0),
                           {(1, 2): (9.0, 9.5), (int(tris[0][0]), int(tris[0][1])): (4.0, 4.5)})
        for k in range(5):
            a, b = rng.choice(n, 2)
            g.add_strut((a, b), 1.5 + k, 0.25)
            o.add_strut(int(a), int(b), 1.5 + k, 0.25)
            g.add_pin(int(a))
            o.add_pin(int(a))
        compare_mesh(g, o)


def test_strip_and_triangle_like_reference_tests():
    # springy_mesh.rs:696-719 and :773-866 through the product library
    pos = [[0, 0, 0], [1, 0, 0], [1, 0, 1], [-1, 1, 0], [0, 1, 0], [1, 1, 0]]
    idx = [0, 4, 3, 0, 1, 4, 1, 5, 4, 1, 2, 5]
    g = fa.SpringyMesh(pos, idx, 1.0, 2.0, 3.0, (4.0, 5.0))
    v01, kcl, fc = g.struts()
    assert v01.tolist() == [[0, 4], [3, 4], [0, 3], [0, 1], [1, 4], [1, 5], [4, 5], [1, 2], [2, 5]]
    has, kcr = g.torsion()
    assert has.tolist() == [True, False, False, False, True, True, False, False, False]
    p, i = g.get_vertices()
    assert p.tolist() == [[float(x) for x in q] for q in pos] and i.tolist() == idx
    t = fa.SpringyMesh([[0, 0, 0], [2, 0, 0], [0, 1, 0]], [0, 1, 2], 1.0, 2.0, 3.0, (4.0, 5.0))
    assert t.get_points()[0, 0] == np.float32(1.0) / np.float32(3.0)
    v01, kcl, _ = t.struts()
    assert kcl[0].tolist() == [1.0, 1.5, 2.0]


def test_errors_replace_panics():
    with pytest.raises(ValueError, match="more than two"):  # springy_mesh.rs:339
        fa.SpringyMesh([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1]], [0, 1, 2, 1, 0, 3, 0, 1, 4],
                       1.0, 1.0, 1.0)
    with pytest.raises(IndexError):
        fa.SpringyMesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [0, 1, 5], 1.0, 1.0, 1.0)
    m = fa.SpringyMesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [0, 1, 2], 1.0, 1.0, 1.0)
    with pytest.raises(IndexError):
        m.add_strut((0, 7), 1.0, 1.0)
    with pytest.raises(ValueError):
        m.set_points(np.zeros((4, 10), np.float32))
    with pytest.raises((ValueError, IndexError)):
        fa.CollidableMesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [0, 0, 1])  # collidable_mesh.rs:83-86
    with pytest.raises(ValueError):
        fa.Cloth(1, 5, 0.1, (0, 0, 0), 1.0, 1, 1, 1, 1, 1, 1, [])


def test_obstacle_counts_match_oracle():
    for p, i in [wl.floor_obstacle(), wl.sphere_obstacle(2.0, (0, 0, 0)), wl.sphere_obstacle(1.0, (1, 2, 3), 3)]:
        g = fa.CollidableMesh(p, i)
        o = orc.Obstacle(p, i)
        assert g.counts() == o.counts()


def test_surrogate_config_default():
    c = _lib.FpSurrogateConfig()
    assert fa.lib().fp_surrogate_config_default(c) == 0
    H = np.float32(np.float32(1.0 / 600.0) * np.float32(8.0))
    assert c.dt == np.float32(1.0 / 600.0) and c.ratio == 8 and c.strain_scale == 10.0
    assert c.out_scale == pytest.approx(float(np.float32(0.005) / (H * H)), rel=1e-6)
    assert c.base_struts == 0 and c.strain_limit == np.float32(0.02) and c.energy_limit == np.float32(0.002)
    d = fa.SurrogateConfig()
    assert (d.strain_limit, d.energy_limit) == (0.02, 0.002)
    assert fa.SurrogateConfig().to_c().out_scale == pytest.approx(c.out_scale, rel=1e-6)



_ASAN_SCRIPT = r"""
import sys
import numpy as np
import feriphys_amd._lib as L
L.LIB_PATH = sys.argv[1]
import feriphys_amd as fa
from feriphys_amd import workloads as wl
cl = fa.Cloth(48, 40, 0.05, (0.0, 0.0, 0.0), 0.01, 500.0, 2.0, 300.0, 1.0, 50.0, 0.5, [0, 39])
m = cl.mesh
v01, kcl, faces = m.struts()
assert m.counts()[1] == 48 * 39 + 47 * 40 + 47 * 39 + 2 * 46 * 38  # tensile, one shear per quad, bends
rec = m.get_points()
m.set_points(rec)
pos, idx = wl.icosphere(3)
s = fa.SpringyMesh(pos, idx, 1.0, 10.0, 0.5)
for a, b in [(0, 5), (3, 9)]:
    s.add_strut((a, b), 2.0, 0.1)
s.add_pin(2)
s.struts(); s.faces(); s.pins(); s.get_vertices()
o = fa.CollidableMesh(*wl.icosphere(2))
o.counts()
print("ok")
"""


def test_host_code_under_address_sanitizer(tmp_path):
    """The library's host code (constructors, tables, readback of host meshes, handles and their
    release) built with AddressSanitizer (make asan; device code unchanged) and driven by the C and
    C++ caller programs, with leak detection, and by the Python mirror with the sanitizer runtime
    preloaded: no invalid access, double free or leak is reported."""
    csrc = os.path.join(ROOT, "feriphys_amd", "csrc")
    subprocess.run(["make", "-s", "-j8", "-C", csrc, "asan"], check=True, capture_output=True)
    lib = os.path.join(ROOT, "_build", "asan", "libferiphys_gpu.so")
    libdir = os.path.dirname(lib)
    llvm = "/opt/rocm/llvm/bin"
    rt = subprocess.run([os.path.join(llvm, "clang"), "-print-file-name=libclang_rt.asan-x86_64.so"],
                        capture_output=True, text=True, check=True).stdout.strip()
    assert os.path.exists(rt), rt
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    for src, cc, std in (("abi_c_check.c", "clang", "-std=c99"), ("mirror_check.cpp", "clang++", "-std=c++17")):
        exe = str(tmp_path / src.split(".")[0])
        subprocess.run([os.path.join(llvm, cc), std, "-g", "-fsanitize=address", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", src), "-L", libdir, "-lferiphys_gpu",
                        "-Wl,-rpath," + libdir, "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0 and r.stdout.strip() == "ok", (src, r.stdout[-2000:], r.stderr[-4000:])
    # the interpreter itself is not instrumented: leaks off, the runtime preloaded
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:verify_asan_link_order=0", LD_PRELOAD=rt,
               PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-c", _ASAN_SCRIPT, lib], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout[-2000:], r.stderr[-4000:])


def test_add_struts_is_the_add_strut_loop():
    """fp_mesh_add_struts appends the same struts as add_strut called once per pair (order, rest
    lengths, constants), and on a bad index appends nothing."""
    v, f = wl.icosphere_fast(2)
    pairs = np.array([[0, 5], [7, 3], [11, 40], [2, 2]], np.uint64)
    a = fa.SpringyMesh(v, f, 1.0, 10.0, 0.5)
    b = fa.SpringyMesh(v, f, 1.0, 10.0, 0.5)
    for p in pairs:
        a.add_strut(p, 3.0, 0.25)
    b.add_struts(pairs, 3.0, 0.25)
    sa, sb = a.struts(), b.struts()
    for x, y in zip(sa, sb):
        assert (np.asarray(x) == np.asarray(y)).all()
    n0 = b.counts()[1]
    with pytest.raises(Exception):
        b.add_struts(np.array([[0, 1], [0, len(v)]], np.uint64), 1.0, 1.0)
    assert b.counts()[1] == n0
