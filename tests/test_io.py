"""Restart primitives and raw Tecplot files written and read by libmoons_amd's
host code, against files the reference's own routines wrote
(tests/golden/io, made by tests/golden/make_io_golden.py with the flang-built
reference) and, where oracle/_ref is built, against those routines directly.

The bar: every double reads back bit for bit in both directions; headers,
zone records, file names and the point order are the reference's; reals are
printed with the fewest digits that read back (Python's repr gives the same
digits) in the F/E layout that flang's list-directed output uses.
"""
import os

import numpy as np
import pytest

import moons_amd as M
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
RAW_CASES = [("CC", 1), ("CC", 0), ("Fx", 0), ("Ey", 1), ("N", 0)]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def item(tmp_path, x):
    """The list-directed item the library writes for the double x."""
    path = str(tmp_path / "one.txt")
    M.gf_write(path, np.full((1, 1, 1), x))
    return open(path).read().splitlines()[2].strip()


# Items flang's list-directed output gives for these doubles (the compiler the
# reference is built with here); each is also the shortest round-trip form.
FLANG_ITEMS = [
    (0.0, "0."), (-0.0, "-0."), (1.0, "1."), (-0.1, "-.1"), (1.0 / 3.0, ".3333333333333333"),
    (1e-300, "1.E-300"), (6.02e23, "6.02E+23"), (-2.5e-7, "-2.5E-07"), (123456789.125, "123456789.125"),
    (1e14, "100000000000000."), (1e15, "1.E+15"), (0.095, ".095"), (0.0949999999, "9.49999999E-02"),
    (0.09500000001, ".09500000001"), (949999999999999.9, "949999999999999.9"), (9.5e14, "9.5E+14"),
    (5e-324, "5.E-324"), (1.7976931348623157e308, "1.7976931348623157E+308"),
    (2.2250738585072014e-308, "2.2250738585072014E-308"), (0.01, "1.E-02"), (0.15000000000000002, ".15000000000000002"),
    (float("inf"), "Inf"), (-float("inf"), "-Inf"), (float("nan"), "NaN"),
]


@pytest.mark.parametrize("x,expect", FLANG_ITEMS)
def test_real_items_match_flang(tmp_path, x, expect):
    assert item(tmp_path, x) == expect


def random_doubles(n, seed):
    rng = np.random.default_rng(seed)
    parts = [
        rng.standard_normal(n) * 10.0 ** rng.integers(-40, 40, n),
        2.0 ** rng.integers(-1074, 1024, n // 4).astype(float),
        np.frombuffer(rng.integers(0, 2**63 - 1, n, dtype=np.int64).tobytes(), dtype=np.float64),
        np.nextafter(10.0 ** rng.integers(-20, 20, n // 4), np.inf),
        np.nextafter(10.0 ** rng.integers(-20, 20, n // 4), 0.0),
    ]
    a = np.concatenate(parts)
    a = a[np.isfinite(a)]
    return a * np.where(rng.random(a.size) < 0.5, -1.0, 1.0)


def digits(tok):
    mant = tok.lstrip("-").upper().split("E")[0].replace(".", "")
    return mant.lstrip("0").rstrip("0") or "0"


def test_shortest_digits_and_bit_exact_read_back(tmp_path):
    a = random_doubles(40000, 3)
    path = str(tmp_path / "r.txt")
    M.gf_write(path, a.reshape(-1, 1, 1))
    toks = open(path).read().split("\n", 2)[2].split("s     =")[0].split()
    assert len(toks) == a.size
    for x, t in zip(a[:20000], toks[:20000]):
        assert float(t) == x
        assert digits(t) == digits(repr(abs(float(x)))), (repr(x), t)
    back = M.gf_read(path).ravel(order="F")
    np.testing.assert_array_equal(bits(back), bits(a))


def test_records_are_filled_greedily_to_79_columns(tmp_path):
    a = random_doubles(3000, 4)
    path = str(tmp_path / "w.txt")
    M.gf_write(path, a.reshape(-1, 1, 1))
    lines = open(path).read().splitlines()
    vals = lines[2:lines.index(" s     = ")]
    assert all(len(l) <= 79 and l.startswith(" ") for l in lines)
    for prev, nxt in zip(vals, vals[1:]):
        assert len(prev) + 1 + len(nxt.split()[0]) > 79


def test_gf_reads_the_reference_file():
    f = np.load(os.path.join(GOLD, "gf_input.npy"))
    g = M.gf_read(os.path.join(GOLD, "gf_reference.txt"))
    assert g.shape == f.shape
    np.testing.assert_array_equal(bits(g), bits(f))


def same_tokens(a, b):
    ta, tb = a.split(), b.split()
    if len(ta) != len(tb):
        return False
    for x, y in zip(ta, tb):
        try:
            fx, fy = float(x), float(y)
            if fx != fy and not (np.isnan(fx) and np.isnan(fy)):
                return False
        except ValueError:
            if x != y:
                return False
    return True


def test_gf_write_matches_the_reference_file(tmp_path):
    f = np.load(os.path.join(GOLD, "gf_input.npy"))
    path = str(tmp_path / "gf.txt")
    M.gf_write(path, f)
    ours, ref = open(path).read(), open(os.path.join(GOLD, "gf_reference.txt")).read()
    assert same_tokens(ours, ref)
    lo, lr = ours.splitlines(), ref.splitlines()
    assert lo[:2] == lr[:2] and lo[-4:] == lr[-4:]  # shape, markers, s, s_1D


def raw_case(name, pad):
    nodes = np.load(os.path.join(GOLD, "raw_nodes.npz"))
    hn = [nodes["hx"], nodes["hy"], nodes["hz"]]
    f = np.load(os.path.join(GOLD, f"raw_input_{name}.npy"))
    d = os.path.join(GOLD, f"raw_{name}_pad{pad}")
    (fname,) = os.listdir(d)
    return hn, f, d + "/", fname


@pytest.mark.parametrize("name,pad", RAW_CASES)
def test_raw_reads_the_reference_file(name, pad):
    hn, f, d, _ = raw_case(name, pad)
    g = M.raw_read(np.full(f.shape, -7.0), name, d, "U", pad)
    inner = tuple(slice(pad, s - pad) for s in f.shape)
    np.testing.assert_array_equal(bits(g[inner]), bits(f[inner]))
    if pad:
        mask = np.ones(f.shape, bool)
        mask[inner] = False
        assert np.all(g[mask] == -7.0)  # pad cells keep their values


@pytest.mark.parametrize("name,pad", RAW_CASES)
def test_raw_write_matches_the_reference_file(tmp_path, name, pad):
    hn, f, d, fname = raw_case(name, pad)
    M.raw_write(hn, name, f, str(tmp_path) + "/", "U", pad)
    assert os.listdir(tmp_path) == [fname]
    ours, ref = open(tmp_path / fname).read(), open(d + fname).read()
    assert ours.splitlines()[:3] == ref.splitlines()[:3]  # title, variables, zone
    assert same_tokens(ours, ref)


def test_raw_unsteady_name_and_errors(tmp_path):
    hn, f, _, _ = raw_case("CC", 1)
    M.raw_write(hn, "CC", f, str(tmp_path) + "/", "p", 0, t=0.5)
    assert os.listdir(tmp_path) == ["pc_t=0.50000000000000000000.dat"]
    with pytest.raises(M.MoonsError, match="cannot read"):
        M.raw_read(f, "CC", str(tmp_path) + "/", "missing", 0)
    # a file cut short is an error, not a partial import
    txt = open(tmp_path / "pc_t=0.50000000000000000000.dat").read().splitlines()
    with open(tmp_path / "qc.dat", "w") as fp:
        fp.write("\n".join(txt[:20]) + "\n")
    with pytest.raises(M.MoonsError, match="end of file"):
        M.raw_read(f, "CC", str(tmp_path) + "/", "q", 0)
    with pytest.raises(M.MoonsError, match="cannot read"):
        M.gf_read(str(tmp_path / "none.txt"))


def test_reader_accepts_other_list_directed_text(tmp_path):
    """gfortran's forms: leading zeros, D exponents, commas and r*c repeats."""
    path = tmp_path / "g.txt"
    path.write_text("  2  2  1\n f     = \n   0.10000000000000001     ,2*1.5D+00\n -2.5000000000000000E-007\n"
                    " s     = \n 2 2 1\n s_1D  = \n 4\n")
    np.testing.assert_array_equal(M.gf_read(str(path)).ravel(order="F"), [0.1, 1.5, 1.5, -2.5e-7])


ref_only = pytest.mark.skipif(not O.ref_available(), reason="reference subset (oracle/_ref) not built here")


@ref_only
@pytest.mark.parametrize("name", list(M.DL))
@pytest.mark.parametrize("pad", [0, 1])
def test_raw_both_ways_through_the_reference_routines(tmp_path, name, pad):
    rng = np.random.default_rng(11 + pad)
    hn = [O.uniform_nodes(0.0, 1.0, 5), O.roberts_both_nodes(-1.0, 1.0, 4, 1.1), O.uniform_nodes(-2.0, 0.0, 3)]
    code = M.DL[name]
    shape = tuple(len(h) if (code >> i) & 1 else len(h) - 1 for i, h in enumerate(hn))
    f = rng.standard_normal(shape) * 10.0 ** rng.integers(-30, 30, shape)
    inner = tuple(slice(pad, s - pad) for s in shape)
    ours, ref = str(tmp_path / "o") + "_", str(tmp_path / "r") + "_"
    M.raw_write(hn, name, f, ours, "B", pad)
    O.ref_export_raw(hn, code, f, ref, "B", pad)
    a = [x for x in os.listdir(tmp_path) if x.startswith("o_")]
    assert len(a) == 1 and os.path.exists(ref + a[0][2:])
    assert same_tokens(open(tmp_path / a[0]).read(), open(ref + a[0][2:]).read())
    g0 = np.zeros(shape)
    np.testing.assert_array_equal(bits(O.ref_import_raw(hn, code, g0, ours, "B", pad)[inner]), bits(f[inner]))
    np.testing.assert_array_equal(bits(M.raw_read(g0, name, ref, "B", pad)[inner]), bits(f[inner]))


@ref_only
def test_gf_both_ways_through_the_reference_routines(tmp_path):
    f = random_doubles(6000, 5)[:5 * 7 * 11].reshape(5, 7, 11, order="F")
    M.gf_write(str(tmp_path / "o.txt"), f)
    O.ref_gf_export(str(tmp_path / "r.txt"), f)
    s, v = O.ref_gf_import(str(tmp_path / "o.txt"), f.size)
    assert s == f.shape
    np.testing.assert_array_equal(bits(v), bits(f.ravel(order="F")))
    np.testing.assert_array_equal(bits(M.gf_read(str(tmp_path / "r.txt"))), bits(f))
    assert same_tokens(open(tmp_path / "o.txt").read(), open(tmp_path / "r.txt").read())
