"""Host-side tests of libttsh3d's C ABI (no GPU calls).

Covers: the library loads and exports every entry point declared in
include/ttsh3d.h; input.dat I/O as GetInput (TTSH3D.cu:696-752) and
run_cuda.pl:53-87 define it, including the -8848 sentinel; parameter
validation; the seeded initial conditions; and the conf_<m>.dat format of
ExpoConf (TTSH3D.cu:966-1013) read back the way readDraw.py:205-215 reads it.
"""
import ctypes
import os
import re

import numpy as np
import pandas as pd
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ttsh3d.h")

# run_cuda.pl:7-44 in the input.dat order of run_cuda.pl:54-86
RUN_CUDA_DEFAULTS = [1, 16, 1, 0, 1, 1, 1, 0, 64, 64, 64, 3, 1, 0.002, 0, 1000, 10, 0.01, 0.000005,
                     1, -1, -0.02, 4.366, -0.187, 1, -0.364, 9, 1, 1, 1, 0, 0.0001, -8848]


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ttsh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(ttsh):
    lib = ctypes.CDLL(ttsh.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header(ttsh):
    ttsh.lib()
    bound = set(ttsh.exported_symbols())
    missing = set(declared_symbols()) - bound
    assert not missing, missing


def test_defaults_match_run_cuda(ttsh):
    p = ttsh.default_params()
    vals = [getattr(p, n) for n in ttsh.INPUT_FIELDS]
    assert vals == pytest.approx(RUN_CUDA_DEFAULTS)
    assert p.faithful_halo == 1


def write_perl_style(path, values):
    # run_cuda.pl prints one interpolated Perl number per line
    with open(path, "w") as f:
        for v in values:
            f.write(f"{v}\n")


def test_parse_run_cuda_input(ttsh, tmp_path):
    vals = list(RUN_CUDA_DEFAULTS)
    vals[8:11] = [128, 128, 128]
    vals[20] = -2  # cA
    path = tmp_path / "input.dat"
    write_perl_style(path, vals)
    p = ttsh.parse_input_dat(str(path))
    assert (p.Nx, p.Ny, p.Nz) == (128, 128, 128)
    assert p.cA == -2 and p.dtMax == 5e-6 and p.lambda0 == 9 and p.ReadSuccess == -8848
    ttsh.validate(p)


def test_input_roundtrip(ttsh, tmp_path):
    p = ttsh.default_params(N=96, cA=-3.25, La=1.4, eta=-0.2, dt0=0.005, Ts=12.5)
    path = str(tmp_path / "input.dat")
    ttsh.write_input_dat(path, p)
    lines = open(path).read().split("\n")
    assert len([x for x in lines if x]) == 33
    q = ttsh.parse_input_dat(path)
    for n in ttsh.INPUT_FIELDS:
        assert getattr(q, n) == getattr(p, n), n


def test_sentinel_failure(ttsh, tmp_path):
    vals = list(RUN_CUDA_DEFAULTS)
    vals[-1] = 1234
    path = tmp_path / "input.dat"
    write_perl_style(path, vals)
    with pytest.raises(ttsh.TTSHError) as e:
        ttsh.parse_input_dat(str(path))
    assert e.value.code == ttsh.ESENTINEL


def test_short_or_missing_input(ttsh, tmp_path):
    path = tmp_path / "input.dat"
    write_perl_style(path, RUN_CUDA_DEFAULTS[:20])
    with pytest.raises(ttsh.TTSHError) as e:
        ttsh.parse_input_dat(str(path))
    assert e.value.code == ttsh.EIO
    with pytest.raises(ttsh.TTSHError) as e:
        ttsh.parse_input_dat(str(tmp_path / "nope.dat"))
    assert e.value.code == ttsh.EIO


@pytest.mark.parametrize("kw,code", [
    (dict(TimeScheme=3), "EUNSUPPORTED"),
    (dict(Nx=64, Ny=64, Nz=32), "EUNSUPPORTED"),
    (dict(N=54), "EUNSUPPORTED"),
    (dict(N=70), "EUNSUPPORTED"),
    (dict(N=45), "EUNSUPPORTED"),
    (dict(N=250), "EUNSUPPORTED"),
    (dict(N=16), "EUNSUPPORTED"),
    (dict(N=512), "EUNSUPPORTED"),
    (dict(Nb=0), "EINVAL"),
    (dict(dt0=0.0), "EINVAL"),
    (dict(ReadSuccess=0), "ESENTINEL"),
])
def test_validation_rejects(ttsh, kw, code):
    p = ttsh.default_params(**kw)
    with pytest.raises(ttsh.TTSHError) as e:
        ttsh.validate(p)
    assert e.value.code == getattr(ttsh, code)


@pytest.mark.parametrize("N", [32, 48, 64, 96, 128, 144, 192, 256])
def test_validation_accepts(ttsh, N):
    ttsh.validate(ttsh.default_params(N=N))


def test_init_droplet_random_unit(ttsh):
    N, R, xi = 32, 10.0, 2.0
    phi, P, v = ttsh.init_droplet(N, R, xi, 0, seed=5)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    assert np.abs(phi - 0.5 * (1 - np.tanh((r - R) / xi))).max() < 1e-14
    inside = phi > 0.5
    n = np.sqrt((P ** 2).sum(0))
    assert np.abs(n[inside] - 1).max() < 1e-12
    assert np.abs(P[:, ~inside]).max() == 0.0
    assert np.abs(v).max() == 0.0
    # isotropic: the mean of each component is small, the second moments equal
    m = P[:, inside].mean(1)
    assert np.abs(m).max() < 0.06
    sec = (P[:, inside] ** 2).mean(1)
    assert np.abs(sec - 1 / 3).max() < 0.03
    # reproducible; different seeds differ
    phi2, P2, _ = ttsh.init_droplet(N, R, xi, 0, seed=5)
    assert np.array_equal(P, P2)
    _, P3, _ = ttsh.init_droplet(N, R, xi, 0, seed=6)
    assert not np.array_equal(P, P3)


def test_init_droplet_aster(ttsh):
    N = 32
    phi, P, v = ttsh.init_droplet(N, 10.0, 2.0, 1, seed=3)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    d = np.stack([x - N / 2, y - N / 2, z - N / 2])
    r = np.sqrt((d ** 2).sum(0))
    inside = (phi > 0.5) & (r > 0)
    cosang = (P * d).sum(0)[inside] / r[inside]
    assert np.median(cosang) > 0.98
    assert np.abs(np.sqrt((P ** 2).sum(0))[inside] - 1).max() < 1e-12


def test_init_reference(ttsh):
    p = ttsh.default_params(N=32, init0=1e-4)
    phi, P, v, vs, ps = ttsh.init_reference(p, seed=1)
    N = 32
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    assert np.abs(phi - (1 - np.tanh((r - N / 4) * 2)) / 2).max() < 1e-14
    for f in (P, v, vs, ps):
        assert np.abs(f).max() <= 0.5e-4
        assert np.abs(f).max() > 0.4e-4
    assert not np.array_equal(v, vs)


def test_conf_format(ttsh, tmp_path):
    N = 32
    rng = np.random.default_rng(0)
    phi = rng.random((N, N, N)).astype(np.float32)
    P = rng.standard_normal((3, N, N, N)).astype(np.float32)
    v = (1e-3 * rng.standard_normal((3, N, N, N))).astype(np.float32)
    path = str(tmp_path / "conf_0.dat")
    ttsh.write_conf_arrays(path, phi, P, v)
    with open(path) as f:
        first = f.readline().rstrip("\n")
    cols = first.split(" ")
    assert len(cols) == 7
    assert all(re.fullmatch(r"-?\d+\.\d{15}", c) for c in cols)
    # readDraw.py:205-215: read_csv(sep=" ") then reshape (Ly, Lx, Lz)
    names = ["vx", "vy", "vz", "px", "py", "pz", "phi"]
    data = pd.read_csv(path, sep=" ", names=names)
    assert len(data) == N ** 3
    arrays = {"vx": v[0], "vy": v[1], "vz": v[2], "px": P[0], "py": P[1], "pz": P[2], "phi": phi}
    for k, a in arrays.items():
        got = data[k].to_numpy().reshape((N, N, N))          # [y, x, z]
        want = np.transpose(a, (1, 2, 0)).astype(np.float64)  # [z, y, x] -> [y, x, z]
        assert np.abs(got - want).max() < 1e-15 + 1e-15 * np.abs(want).max()


def test_conf_text_is_printf_exact(ttsh, tmp_path):
    """Every value of a snapshot is the "%.15f" text of the FP32 value as a
    double (TTSH3D.cu:983-998: std::fixed, setprecision(15)), byte for byte:
    random finite bit patterns over the whole exponent range, exact halfway
    cases at the 15th decimal (odd multiples of 2^-16 round half to even),
    signed zeros, subnormals, the 2^40 boundary of the integer path, and
    non-finite values."""
    N = 16
    rng = np.random.default_rng(11)
    n = 7 * N ** 3
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32).copy()
    vals[~np.isfinite(vals)] = 1.5
    special = [0.0, -0.0, 2.0 ** -16, 3 * 2.0 ** -16, -5 * 2.0 ** -16, 2.0 ** -17, 1e-45, -1e-45, 1.17549435e-38,
               0.5e-15, 1.5e-15, 2.5e-15, 1.0, -1.0, 123.456, 2.0 ** 39, 2.0 ** 40 - 2.0 ** 16, 2.0 ** 40, -2.0 ** 41,
               3.4e38, np.inf, -np.inf, np.nan, 0.1, 0.7, 0.999999999, 33554431.0]
    vals[:len(special)] = np.array(special, np.float32)
    # halfway cases k * 2^-16 for odd k, and small multiples of 2^-17..2^-20
    m = 2000
    vals[100:100 + m] = (rng.integers(-(2 ** 20), 2 ** 20, m) * 2 + 1).astype(np.float32) * np.float32(2.0 ** -16)
    vals[3000:3000 + m] = rng.integers(-(2 ** 20), 2 ** 20, m).astype(np.float32) * np.float32(2.0 ** -19)
    cube = vals.reshape(7, N, N, N)
    v, P, phi = cube[0:3], cube[3:6], cube[6]
    path = str(tmp_path / "conf_x.dat")
    ttsh.write_conf_arrays(path, phi, P, v)
    got = open(path).read().split()
    # the row order of ExpoConf: y outer, x, z inner; fields vx vy vz px py pz phi
    order = np.stack([v[0], v[1], v[2], P[0], P[1], P[2], phi])  # [f, z, y, x]
    want = ["%.15f" % float(x) for x in np.transpose(order, (2, 3, 1, 0)).reshape(-1)]
    assert len(got) == len(want)
    bad = [(g, w) for g, w in zip(got, want) if g != w]
    assert not bad, bad[:5]


def test_conf_format_expo_minus_one(ttsh, tmp_path):
    N = 32
    z = np.zeros((N, N, N), np.float32)
    Z = np.zeros((3, N, N, N), np.float32)
    path = str(tmp_path / "conf_m.dat")
    ttsh.write_conf_arrays(path, z, Z, Z, expo_data=-1)
    txt = open(path).read()
    assert "\n" not in txt
    assert txt.count(" ") == 2 * N ** 3


def test_droplet_bytes(ttsh):
    # 20 FP32 fields, the complex projection buffer (2) and the edge-column
    # buffer (1), plus the per-item sums
    b = ttsh.droplet_bytes(128)
    assert 23 * 4 * 128 ** 3 <= b < 24 * 4 * 128 ** 3


def test_cli_rejects_bad_sentinel(tmp_path):
    import subprocess
    exe = os.path.join(ROOT, "bacterial-droplet_amd", "bin", "ttsh3d")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    d = tmp_path / "data_7"
    d.mkdir()
    vals = list(RUN_CUDA_DEFAULTS)
    vals[-1] = 0
    write_perl_style(d / "input.dat", vals)
    r = subprocess.run([exe, "7"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    assert "Error while reading the input file!" in r.stdout


def test_tuning_table_batch_line(ttsh, tmp_path):
    """The tuning table's per-size lines and its device-filling batch line
    ("batch rhs_variant rhs_tz corr_tz") round-trip through the library."""
    from ttsh3d import autotune
    path = str(tmp_path / "tuned.txt")
    autotune.write_table({64: (4, 8, 4), 128: (5, 16, 16)}, path, (1, 48, 16))
    assert autotune.read_table(path) == ({64: (4, 8, 4), 128: (5, 16, 16)}, (1, 48, 16))
    try:
        ttsh.tuning_reload(path)
        assert ttsh.tuning_lookup(128) == (5, 16, 16)
        assert ttsh.tuning_lookup(96) is None
        assert ttsh.tuning_lookup_batch() == (1, 48, 16)
        autotune.write_table({64: (4, 8, 4)}, path)
        ttsh.tuning_reload(path)
        assert ttsh.tuning_lookup_batch() is None
    finally:
        ttsh.tuning_reload(None)


PLANNED = [32, 36, 40, 48, 50, 60, 64, 72, 80, 90, 96, 100, 120, 128, 144, 150, 160, 180, 192, 200, 240, 256]


def test_planned_sizes_are_the_even_smooth_sizes_without_27_or_125():
    """Every even N in [32, 256] with prime factors 2, 3, 5 splits as R1 x R2
    over the register DFT lengths {2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20}
    unless 27 or 125 divides it; those are the planned sizes."""
    def smooth(n):
        for q in (2, 3, 5):
            while n % q == 0:
                n //= q
        return n == 1
    lengths = {2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20}
    even = [n for n in range(32, 257, 2) if smooth(n)]
    assert [n for n in even if n % 27 and n % 125] == PLANNED
    assert [n for n in even if any(n % a == 0 and n // a in lengths for a in lengths)] == PLANNED


def test_shipped_tuning_table_covers_every_planned_size(ttsh):
    """tuned.txt beside the library holds an entry for every planned size, each
    with a K_rhs tile that fits the size, and the batch line; the library
    reads the same entries back."""
    from ttsh3d import autotune
    entries, batch = autotune.read_table(autotune.DEFAULT_TABLE)
    assert sorted(entries) == PLANNED
    assert batch is not None
    for N, (v, rtz, ptz) in entries.items():
        tx, ty = autotune.RHS_VARIANTS[v]
        assert N >= tx + 2 and N >= ty + 2 and 0 < rtz <= N and 0 < ptz <= N, (N, v, rtz, ptz)
    ttsh.tuning_reload(None)
    for N in (36, 128, 256):
        assert ttsh.tuning_lookup(N) == entries[N]


@pytest.mark.parametrize("N", [54, 108, 162, 216, 250])
def test_sizes_without_a_plan_are_refused(ttsh, N):
    p = ttsh.default_params(N=N)
    assert ttsh.lib().ttsh_validate_params(ctypes.byref(p)) == ttsh.EUNSUPPORTED


@pytest.mark.parametrize("N", PLANNED)
def test_every_planned_size_validates(ttsh, N):
    """The sizes with a compile-time FFT plan."""
    p = ttsh.default_params(N=N)
    assert ttsh.lib().ttsh_validate_params(ctypes.byref(p)) == 0
