"""CPU-only checks of librsf.so: it loads, exports every symbol include/rsf.h
declares, and its host-side parts (taps, phantom generator, signed distance,
CLI argument and .npy handling) behave; no GPU compute is called here."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import rsf_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rsf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rsf_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 38
    L = rsf_amd.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding knows every one of them
    assert sorted(set(names) - set(rsf_amd.EXPORTED)) == []


def test_cli_library_exports_its_entry_point():
    import ctypes

    text = open(os.path.join(ROOT, "include", "rsf_cli.h")).read()
    assert "int rsf_cli_main(int argc, char** argv);" in text
    rsf_amd.lib()
    L = ctypes.CDLL(os.path.join(rsf_amd.PKG_DIR, "librsf_cli.so"))
    assert hasattr(L, "rsf_cli_main")


def test_abi_version_and_status_strings():
    L = rsf_amd.lib()
    assert L.rsf_abi_version() == 1
    assert L.rsf_status_string(0) == b"ok"
    assert L.rsf_status_string(1) == b"invalid argument"


def test_reference_defaults():
    p = rsf_amd.reference_defaults()  # CUDA_RSF/main.cpp:91-97
    assert (p.sigma, p.iterations, p.mu) == (5.0, 30, 1.0)
    assert abs(p.dt - 0.1) < 1e-7 and abs(p.lambda1 - 0.1) < 1e-7 and abs(p.nu - 1e-4) < 1e-10
    assert p.sigma2 == 0.0 and abs(p.epsilon - 0.2) < 1e-7
    assert p.mode == rsf_amd.MODE_LITERAL   # the reference's arithmetic unless asked otherwise
    assert rsf_amd.make_params().mode == rsf_amd.MODE_LITERAL


@pytest.mark.parametrize("sigma", [1.0, 2.5, 3.0, 5.0])
def test_taps_match_oracle(sigma):
    g, P, m = rsf_amd.taps(sigma)
    go, Po, mo = oracle.taps(sigma)
    assert (P, m) == (Po, mo)
    assert np.array_equal(g, go)


@pytest.mark.parametrize("sigma", [2.0, 4.0, 0.1, -1.0])
def test_taps_reject(sigma):
    with pytest.raises(rsf_amd.RSFError):
        rsf_amd.taps(sigma)


def test_context_without_device_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(rsf_amd.RSFError) as e:
        rsf_amd.Context(0)
    assert e.value.status == rsf_amd.ERR_DEVICE


# ----------------------------------------------------------------- generator
def test_synth_tubes_deterministic_and_shaped():
    a, ma = rsf_amd.synth_tubes((40, 44, 48), tubes=6, seed=3)
    b, mb = rsf_amd.synth_tubes((40, 44, 48), tubes=6, seed=3)
    c, _ = rsf_amd.synth_tubes((40, 44, 48), tubes=6, seed=4)
    assert np.array_equal(a, b) and np.array_equal(ma, mb) and not np.array_equal(a, c)
    assert a.dtype == np.uint8 and set(np.unique(ma)) <= {0, 1} and 0 < ma.mean() < 0.5
    # no noise, no bias: exactly 200 in the tubes and 30 outside
    d, md = rsf_amd.synth_tubes((40, 44, 48), tubes=6, seed=3, noise=0.0, bias=0.0)
    assert np.array_equal(md, ma)
    assert np.all(d[md == 1] == 200) and np.all(d[md == 0] == 30)
    # the bias is a linear x ramp of +-20
    e, _ = rsf_amd.synth_tubes((8, 8, 41), tubes=0, noise=0.0, bias=20.0)
    assert e[0, 0, 0] == 10 and e[0, 0, -1] == 50 and e[0, 0, 20] == 30
    # noise sigma ~ 15 on the background
    f, mf = rsf_amd.synth_tubes((64, 64, 64), tubes=0, noise=15.0, bias=0.0, seed=9)
    assert abs(float(f.std()) - 15.0) < 0.3 and abs(float(f.mean()) - 30.0) < 0.2


def test_initial_phi_is_exact_signed_distance():
    from scipy import ndimage

    img, _ = rsf_amd.synth_tubes((36, 40, 44), tubes=5, seed=5)
    phi = rsf_amd.initial_phi(img, 1.0, 110.0)
    m = ndimage.gaussian_filter(img.astype(np.float64), 1.0, mode="nearest", truncate=3.0) > 110
    ref = np.where(m, -(ndimage.distance_transform_edt(m) - 0.5), ndimage.distance_transform_edt(~m) - 0.5)
    assert np.abs(phi - ref).max() < 1e-4
    assert np.array_equal(phi < 0, m)


def test_sdf_from_seed_sign_and_distance():
    from scipy import ndimage

    rng = np.random.default_rng(1)
    seed = np.zeros((20, 22, 24), np.float32)
    seed[5:12, 6:15, 4:20] = 255.0
    seed[15:18, 2:5, 2:5] = 14.0     # just above the shift of 13: inside
    seed[1:3, 18:20, 20:22] = 13.0   # exactly 13: outside (seed - 13 > 0 is false)
    phi = rsf_amd.sdf_from_seed(seed, 13.0)
    region = seed - 13.0 > 0
    ref = np.where(region, ndimage.distance_transform_edt(region) - 0.5,
                   -(ndimage.distance_transform_edt(~region) - 0.5))
    assert np.abs(phi - ref).max() < 1e-4
    assert np.all(phi[region] > 0) and np.all(phi[~region] < 0)
    _ = rng


# ----------------------------------------------------------------------- CLI
def _cli(*args, cwd=None):
    return subprocess.run([rsf_amd.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_help_exits_1_like_the_reference():
    r = _cli("--help")
    assert r.returncode == 1  # CUDA_RSF/main.cpp:109-112
    for opt in ("--sigma", "--t", "--dt", "--meu", "--e", "--v", "--sigma2", "--binary", "--image", "--ouput"):
        assert opt in r.stdout


def test_cli_rejects_unknown_and_ambiguous_options():
    assert _cli("--nope", "1").returncode != 0
    r = _cli("--sig", "3")  # --sigma / --sigma2
    assert r.returncode != 0 and "ambiguous" in r.stderr
    assert _cli("-t", "3").returncode != 0  # short options are disabled (main.cpp:104-106)


def _save_raw_npy(path, header, payload):
    h = header.encode()
    total = 10 + len(h) + 1
    h += b" " * ((64 - total % 64) % 64) + b"\n"
    with open(path, "wb") as f:
        f.write(b"\x93NUMPY\x01\x00" + len(h).to_bytes(2, "little") + h + payload)


def test_cli_input_errors(tmp_path):
    img = tmp_path / "img.npy"
    seed = tmp_path / "seed.npy"
    np.save(img, np.zeros((24, 24, 24), np.uint8))
    np.save(seed, np.zeros((24, 24, 25), np.uint8))
    r = _cli("--image", str(img), "--binary", str(seed), "--t", "1")
    assert r.returncode != 0 and "shapes differ" in r.stderr
    r = _cli("--image", str(tmp_path / "missing.npy"), "--binary", str(seed))
    assert r.returncode != 0 and "cannot open" in r.stderr
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"not a numpy file at all")
    assert "not a .npy" in _cli("--image", str(bad), "--binary", str(seed)).stderr
    trunc = tmp_path / "trunc.npy"
    _save_raw_npy(trunc, "{'descr': '|u1', 'fortran_order': False, 'shape': (24, 24, 24), }", b"\x00" * 100)
    assert "truncated" in _cli("--image", str(trunc), "--binary", str(seed)).stderr
    fort = tmp_path / "fort.npy"
    np.save(fort, np.asfortranarray(np.zeros((24, 24, 24), np.uint8)))
    assert "Fortran" in _cli("--image", str(fort), "--binary", str(seed)).stderr
    cplx = tmp_path / "cplx.npy"
    np.save(cplx, np.zeros((24, 24, 24), np.complex64))
    assert "unsupported dtype" in _cli("--image", str(cplx), "--binary", str(seed)).stderr
    flat = tmp_path / "flat.npy"
    np.save(flat, np.zeros((24, 24), np.uint8))
    assert "3-D" in _cli("--image", str(flat), "--binary", str(seed)).stderr
    assert "--binary" in _cli("--image", str(img)).stderr
    assert _cli("--image", str(img), "--binary", str(img), "--t", "2.5").returncode != 0
    assert _cli("--image", str(img), "--binary", str(img), "--sigma", "abc").returncode != 0


def test_cli_block_and_batch_argument_errors(tmp_path):
    img = tmp_path / "img.npy"
    np.save(img, np.zeros((24, 24, 24), np.uint8))
    for bad in ("0", "32,32", "a", "32,32,32,32", "-4"):
        r = _cli("--image", str(img), "--binary", str(img), "--block", bad)
        assert r.returncode != 0 and "--block" in r.stderr, bad
    r = _cli("--image", str(img), "--binary", str(img), "--block", "16", "--overlap", "2.5")
    assert r.returncode != 0 and "--overlap" in r.stderr
    assert "cannot open manifest" in _cli("--batch", str(tmp_path / "none.txt")).stderr
    man = tmp_path / "m.txt"
    man.write_text("# blocks\n\nimg.npy img.npy\n")
    r = _cli("--batch", str(man))
    assert r.returncode != 0 and "m.txt:3" in r.stderr
    man.write_text("# nothing\n")
    assert "lists no blocks" in _cli("--batch", str(man)).stderr
    man.write_text("img.npy missing.npy out.npy\n")   # paths relative to the manifest
    assert "missing.npy" in _cli("--batch", str(man)).stderr
    assert _cli("--batch", str(man), "--image", str(img)).returncode != 0


def test_npy_info_reads_headers(tmp_path):
    cases = {"u8": np.zeros((3, 4, 5), np.uint8), "f4": np.zeros((2, 3, 4), np.float32),
             "f8": np.zeros((6, 1, 2), np.float64), "i2": np.zeros((7,), np.int16), "b1": np.zeros((2, 2, 2), bool)}
    for name, a in cases.items():
        path = tmp_path / f"{name}.npy"
        np.save(path, a)
        shape, u8 = rsf_amd.npy_info(str(path))
        assert shape == a.shape and u8 == (a.dtype in (np.uint8, np.bool_)), name
    trunc = tmp_path / "trunc.npy"
    _save_raw_npy(trunc, "{'descr': '<f4', 'fortran_order': False, 'shape': (4, 4, 4), }", b"\x00" * 12)
    with pytest.raises(rsf_amd.RSFError, match="truncated"):
        rsf_amd.npy_info(str(trunc))
    (tmp_path / "bad.npy").write_bytes(b"nope")
    with pytest.raises(rsf_amd.RSFError, match="not a .npy"):
        rsf_amd.npy_info(str(tmp_path / "bad.npy"))


def test_variant_query_names_the_tuner_candidates():
    """The variant query lists what rsf_autotune times: split and fused sweep-1
    kernels for P = 21 in both modes (among them the tuner's batched and
    large-volume choices), a fused sweep-1 kernel for every odd tap count up to
    35, and sweep-2 kernels of both modes; out-of-range queries are errors."""
    v1, v2 = rsf_amd.variants(1), rsf_amd.variants(2)
    names1 = {n for n, _, _ in v1}
    assert {"p21_lit_hi_32x44x8", "p21_lit_hi_32x60r15", "p21_fast_split_64x44w4", "p21_lit_strip_64w4",
            "p21_lit_strip_64w3", "p21_lit_strip_64z1k", "p21_fast_strip_128w4"} <= names1
    # the defaults (the first P = 21 entry of each mode) are the strip kernels
    assert next(n for n, t, l in v1 if t == 21 and l) == "p21_lit_strip_64z1k"
    assert next(n for n, t, l in v1 if t == 21 and not l) == "p21_fast_strip_128w4"
    for P in range(7, 37, 2):
        for lit in (False, True):
            assert any(t == P and l == lit for _, t, l in v1), (P, lit)
    assert all(t == 0 for _, t, _ in v2)
    assert any(l for _, _, l in v2) and any(not l for _, _, l in v2)
    import ctypes as C

    L = rsf_amd.lib()
    n = C.c_int()
    assert L.rsf_variant_count(3, C.byref(n)) == rsf_amd.ERR_INVALID_ARGUMENT
    assert L.rsf_variant_info(1, len(v1), None, 0, None, None) == rsf_amd.ERR_INVALID_ARGUMENT


def test_device_wrappers_refuse_host_and_mismatched_tensors():
    """The device-pointer wrappers hand tensor addresses to kernels: host
    tensors, numpy arrays, wrong dtypes, non-contiguous views and tensors
    smaller than the blocks are refused before any library call."""
    import torch

    chk = rsf_amd._device_tensor
    with pytest.raises(TypeError):
        chk(np.zeros(8, np.float32), "phi")
    with pytest.raises(ValueError):
        chk(torch.zeros(8), "phi")           # a host tensor

    class FakeDev:
        """Stands in for a device tensor (no HIP device here)."""
        def __init__(self, t):
            self.t, self.is_cuda, self.dtype, self.shape, self.device = t, True, t.dtype, t.shape, "cuda:0"

        def data_ptr(self):
            return 0

        def is_contiguous(self):
            return self.t.is_contiguous()

        def numel(self):
            return self.t.numel()

    chk(FakeDev(torch.zeros(2, 3, 4)), "phi", numel=24, shape=(2, 3, 4))
    with pytest.raises(TypeError):
        chk(FakeDev(torch.zeros(8, dtype=torch.float64)), "phi")
    with pytest.raises(ValueError):
        chk(FakeDev(torch.zeros(4, 6).t()), "phi")         # not contiguous
    with pytest.raises(ValueError):
        chk(FakeDev(torch.zeros(20)), "phi", numel=24)      # smaller than the blocks
    with pytest.raises(ValueError):
        chk(FakeDev(torch.zeros(2, 4, 3)), "phi", shape=(2, 3, 4))
    chk(FakeDev(torch.zeros(8, dtype=torch.uint8)), "image", dtypes=rsf_amd._IMAGE_DTYPES)
    assert rsf_amd._extent([(0, 4, 4, 4), (64, 2, 2, 2)]) == 72


def test_workspace_size_counts_the_split_sweep_buffer():
    """rsf_workspace_size of a plain run: phi ping-pong, transform and force, plus
    the split sweep-1 xy channels (4 literal, 2 fast); fast mode adds blur(I) and
    the float image, and two pass temporaries where no split kernel blurs the image; blocks of 2^30 voxels and more run the
    fused sweep 1 without the xy buffer; tap counts without a fused kernel take the
    unfused pipeline's 22 volumes."""
    import ctypes as C

    L = rsf_amd.lib()

    def ws(mode, w, h, l, sigma=3.0):
        b = C.c_int64()
        rsf_amd.check(L.rsf_workspace_size(C.byref(rsf_amd.make_params(sigma=sigma, mode=mode)), w, h, l,
                                           C.byref(b)))
        return b.value

    n = 64 * 70 * 80
    assert ws(rsf_amd.MODE_LITERAL, 64, 70, 80) == (3 + 4) * 4 * n
    assert ws(rsf_amd.MODE_FAST, 64, 70, 80) == (3 + 2 + 2) * 4 * n
    # no split kernel at this tap count: the per-op image blur's two pass temporaries
    assert ws(rsf_amd.MODE_FAST, 64, 70, 80, sigma=2.5) == (3 + 2) * 4 * n + 2 * 4 * (64 - 16) * 70 * 80
    big = (1024, 1024, 1024)
    assert ws(rsf_amd.MODE_LITERAL, *big) == 3 * 4 * 1024 ** 3
    assert ws(rsf_amd.MODE_LITERAL, 64, 70, 80, sigma=5.3) == 22 * 4 * n
    b = C.c_int64()
    assert L.rsf_workspace_size(C.byref(rsf_amd.make_params()), 10, 70, 80, C.byref(b)) == rsf_amd.ERR_INVALID_ARGUMENT
