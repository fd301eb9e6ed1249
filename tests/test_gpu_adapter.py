"""The C++ host classes (include/rapidnj_amd_distance.hpp) used the way rapidNJ's main uses
KimuraDistance / JCdistance: in-memory rows, caller-owned (bootstrap) rows and the disk-matrix
row stream. Expected matrices are the reference engine's own (tests/golden/)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "adapter_test")
GOLDEN = os.path.join(HERE, "golden")


def _write_input(path, d):
    aln = np.ascontiguousarray(d["aln"], dtype=np.uint8)
    n, L = aln.shape
    if "bits" in d:
        bits, gaps = d["bits"], d["gaps"]
    else:
        from oracle import oracle as O
        bits, gaps, _ = O.pack(aln)
    with open(path, "wb") as f:
        np.array([n, L, bits.shape[1] // 4], dtype=np.uint32).tofile(f)
        aln.tofile(f)
        np.ascontiguousarray(bits, dtype=np.uint32).tofile(f)
        np.ascontiguousarray(gaps, dtype=np.uint32).tofile(f)


def _run(tmp_path, d, model, mode, source, protein=False):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, d)
    extra = ["protein"] if protein else []
    r = subprocess.run([DRIVER, str(inp), str(out), model, mode, source] + extra,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint32)
    n = int(raw[0])
    mat = raw[1:1 + n * n].view(np.float32).reshape(n, n)
    mx = raw[1 + n * n:].view(np.float32)[0]
    return mat, mx


# "reuse" (rows handed in by the caller) exists for KimuraDistance only: the reference's
# JCdistance has no caller-rows constructor
MODES = [(model, mode) for model in ("kim", "jc") for mode in ("mem", "disk", "reuse", "half")
         if not (mode == "reuse" and model == "jc")]


@pytest.mark.parametrize("case", ["lgt70.npz", "c1.npz", "saturated.npz", "edge_L129.npz"])
@pytest.mark.parametrize("model,mode", MODES)
def test_host_classes_match_reference(tmp_path, case, model, mode):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run build()")
    d = dict(np.load(os.path.join(GOLDEN, case), allow_pickle=False))
    ref = d["k2p" if model == "kim" else "jc"]
    mat, mx = _run(tmp_path, d, model, mode, "fast")
    np.testing.assert_allclose(mat, ref, rtol=1e-6, atol=0)
    np.testing.assert_array_equal(np.signbit(mat), np.signbit(ref))
    off = ~np.eye(mat.shape[0], dtype=bool)
    if np.any(ref[off] == np.float32(2) * np.float32(mx)):   # saturated pairs carry 2 * max
        assert mx > 0


def test_character_rows_path(tmp_path):
    """fastdist off: the loader's character rows are packed on the device (uppercase ACGT and
    '-' give the same counts as the bit strings)."""
    d = dict(np.load(os.path.join(GOLDEN, "c1.npz"), allow_pickle=False))
    mat, _ = _run(tmp_path, d, "kim", "mem", "chars")
    np.testing.assert_allclose(mat, d["k2p"], rtol=1e-6, atol=0)


GAP_CHARS = b"-.XxzZbBJj?"


def _protein_loader_data(aln):
    """The arrays the reference's FASTA loader keeps for a protein alignment: resolved characters
    (gap characters as '-', dataloaderFasta.hpp resolveChar), padded with '-' to bitStringsCount
    = L/16 + 8 blocks of 16, four to a u32 (encodeProteinSequence)."""
    n, L = aln.shape
    res = aln.copy()
    res[np.isin(res, np.frombuffer(GAP_CHARS, dtype=np.uint8))] = ord("-")
    bsc = L // 16 + 8
    padded = np.full((n, bsc * 16), ord("-"), dtype=np.uint8)
    padded[:, :L] = res
    bits = padded.view("<u4").reshape(n, bsc * 4)
    return {"aln": res, "bits": bits, "gaps": np.zeros_like(bits)}


@pytest.mark.parametrize("model", ["kim", "jc"])
@pytest.mark.parametrize("mode", ["mem", "disk", "half"])
@pytest.mark.parametrize("source", ["fast", "chars"])
def test_protein_loader(tmp_path, model, mode, source):
    """A protein loader (type PROTEIN) gets the reference's protein distances through the same
    classes: PF15224 from the reference's test data, and an alignment with a multiple-of-128
    length and an all-gap row. With fastdist on ("fast") the reference counts with
    bitDistanceProtein; with fastdist off ("chars", the -f path) with
    simpleDistanceCalculator::calculateHammingDistance, whose valid count differs on every pair
    of p20x128 (L = 128)."""
    g = dict(np.load(os.path.join(GOLDEN, "protein.npz"), allow_pickle=False))
    key = model if source == "fast" else "n" + model
    for case in ("pf15224", "p20x128", "p10x200gap"):
        d = _protein_loader_data(g[f"{case}_aln"])
        mat, _ = _run(tmp_path, d, model, mode, source, protein=True)
        want = g[f"{case}_{key}"]
        np.testing.assert_allclose(mat, want, rtol=1e-6, atol=0)
        np.testing.assert_array_equal(np.signbit(mat), np.signbit(want))


@pytest.mark.parametrize("case", ["mixed", "sim"])
@pytest.mark.parametrize("model,key", [("kim", "k2p"), ("jc", "jc")])
@pytest.mark.parametrize("mode", ["mem", "disk"])
def test_naive_rows_memory_and_disk(tmp_path, case, model, key, mode):
    """fastdist off on mixed-case DNA: the in-memory matrix is the mirrored upper triangle, the
    disk matrix holds every (i, j) counted with i as the row (the reference's two paths, written
    by the reference in tests/golden/naive_rows.npz)."""
    g = dict(np.load(os.path.join(GOLDEN, "naive_rows.npz"), allow_pickle=False))
    mat, _ = _run(tmp_path, {"aln": g[f"{case}_aln"]}, model, mode, "chars")
    want = g[f"{case}_{mode}_{key}"]
    np.testing.assert_allclose(mat, want, rtol=1e-6, atol=0)
    np.testing.assert_array_equal(np.signbit(mat), np.signbit(want))


@pytest.mark.parametrize("case,model", [("c1", "jc"), ("c1", "kim"), ("saturated", "kim"),
                                        ("lgt70", "kim"), ("pf15224", "kim")])
def test_printed_matrix_matches_reference_program(tmp_path, case, model):
    """The reference's `-o m` output (printDistanceMatrix, src/main.cpp:212-221), printed by the
    reference's own program from its own matrix (tests/golden/matrix_text.npz), equals the text
    rapidnj_amd::printDistanceMatrix writes for the rows the GPU calculator hands over, byte for
    byte: 6 fixed decimals absorb the last-ulp differences the 1e-6 tolerance allows."""
    g = dict(np.load(os.path.join(GOLDEN, "matrix_text.npz"), allow_pickle=False))
    aln = g[f"{case}_aln"]
    protein = case == "pf15224"
    d = _protein_loader_data(aln) if protein else {"aln": aln}
    inp, out = tmp_path / "in.bin", tmp_path / "out.txt"
    _write_input(inp, d)
    r = subprocess.run([DRIVER, str(inp), str(out), model, "print", "fast"] +
                       (["protein"] if protein else []), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    want = g[f"{case}_{model}_text"].tobytes()
    got = out.read_bytes()
    assert got.split(b"\n")[0] == want.split(b"\n")[0]
    assert got == want
