"""End-to-end drop-in test of the kestrel CLI on the GPU: VCF in, TSV out
(cli.rs:129-181 columns), against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from kestrel_amd import segment_bounds, synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KESTREL = os.path.join(ROOT, "kestrel_amd", "kestrel")
KIN = np.array([1, 0, .5, 0, .5, 0, .5, .25, 0])
COLS = ["sample1", "sample2", "kinship"] + [f"jacquard{i}" for i in range(1, 10)] + ["iterations", "objective",
                                                                                        "condition"]


def log10(a):
    """libm log10, as Rust's f64::log10 and the CLI use (numpy's may differ by an ulp)."""
    return np.array([math.log10(v) if v > 0 else (np.nan if np.isnan(v) else -np.inf) for v in a])


def run(vcf, out, *args):
    r = subprocess.run([KESTREL, vcf, out, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = open(out).read().rstrip("\n").split("\n")
    assert lines[0].split("\t") == COLS
    rows = [ln.split("\t") for ln in lines[1:]]
    return rows, r.stderr


def as_arrays(rows, names):
    idx = {n: i for i, n in enumerate(names)}
    x = np.array([idx[r[0]] for r in rows])
    y = np.array([idx[r[1]] for r in rows])
    kin = np.array([float(r[2]) for r in rows])
    jac = np.array([[float(v) for v in r[3:12]] for r in rows])
    it = np.array([int(r[12]) for r in rows])
    obj = np.array([float(r[13]) for r in rows])
    cond = np.array([float(r[14]) for r in rows])
    return x, y, kin, jac, it, obj, cond


def test_cli_gt_cls_bitexact(O, tmp_path):
    print("case: cli -f GT", flush=True)
    c = synth.generate(30, 2500, seed=21, n_family_pairs=6, missing=0.002)
    vcf = str(tmp_path / "c.vcf.gz")
    synth.write_vcf(c, vcf, "GT")
    rows, log = run(vcf, str(tmp_path / "o.tsv"), "-i", "100")
    x, y, kin, jac, it, obj, cond = as_arrays(rows, c.names)
    keep = (c.codes != 3).all(0)
    exp, _ = O.cls_fit_pairs(c.codes[:, keep])
    assert np.array_equal(x, exp["x"]) and np.array_equal(y, exp["y"])
    assert np.array_equal(jac, exp["jacquard"]) and np.array_equal(it, exp["iters"])
    assert np.array_equal(obj, exp["obj"])
    assert np.array_equal(cond, log10(exp["cond"]), equal_nan=True)
    kin_exp = np.zeros(len(exp))
    for k in range(9):
        kin_exp = kin_exp + exp["jacquard"][:, k] * KIN[k]
    assert np.array_equal(kin, kin_exp)
    assert "Wall Time" in log


@pytest.mark.parametrize("af_info", [False, True])
def test_cli_gt_multiallelic_cls_bitexact(O, tmp_path, af_info):
    """-f GT with 2-, 3- and 4-allele records: the reader's buckets, the
    three-bucket CLS fit (coefficients.rs:166-320), TSV text."""
    print(f"case: cli -f GT multi-allelic af_info={af_info}", flush=True)
    idx = synth.multi_genotypes(22, (400, 60, 30), seed=23, missing=0.002)
    rng = np.random.default_rng(4)
    af = [rng.dirichlet(np.ones(A), v) for A, v in zip((2, 3, 4), (400, 60, 30))] if af_info else None
    vcf = str(tmp_path / "m.vcf")
    synth.write_vcf_gt_multi(vcf, idx, af=af)
    rows, log = run(vcf, str(tmp_path / "o.tsv"), *(["-a", "info"] if af_info else []))
    names = [f"S{i:05d}" for i in range(22)]
    x, y, kin, jac, it, obj, cond = as_arrays(rows, names)
    keep = [~(b == 255).any(axis=0) for b in idx]
    afk = None
    if af_info:  # the text the writer put in INFO, read as f32 (vcf.rs:446-482)
        afk = []
        for a, k in zip(af, keep):
            alt = np.array([[float(np.float32(float(f"{v:.7g}"))) for v in row[1:]] for row in a[k]])
            ref = 1.0 - np.clip(np.array([sum(r, 0.0) for r in alt]), 0.0, 1.0)
            afk.append(np.column_stack([ref, alt]))
    exp, _ = O.cls_fit_pairs_multi(*(b[:, k] for b, k in zip(idx, keep)), af=afk)
    assert np.array_equal(x, exp["x"]) and np.array_equal(y, exp["y"])
    assert np.array_equal(jac, exp["jacquard"]) and np.array_equal(it, exp["iters"])
    assert np.array_equal(obj, exp["obj"])
    assert np.array_equal(cond, log10(exp["cond"]), equal_nan=True)
    assert "tri-allelic" in log


@pytest.mark.parametrize("field", ["GL", "PL"])
def test_cli_ml_matches_oracle(O, tmp_path, field):
    print(f"case: cli -f {field}", flush=True)
    c = synth.generate(28, 3000, seed=22, n_family_pairs=6, missing=0.02)
    vcf = str(tmp_path / f"c.{field}.vcf")
    synth.write_vcf(c, vcf, field)
    rows, _ = run(vcf, str(tmp_path / "o.tsv"), "-f", field, "-s", "20")
    x, y, kin, jac, it, obj, cond = as_arrays(rows, c.names)
    assert len(rows) == 20 * 21 // 2 and x.max() < 20
    af = O.allele_freqs(c.codes)  # frequencies use every sample (cli.rs:41-45)
    kept = O.keep_sites(af)
    codes, afk = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    exp = O.fit_pairs(codes, afk, nsel=20, order=O.ORDER_CLASS)  # hard calls: the class-sorted sweep
    assert np.array_equal(jac, exp["jacquard"]) and np.array_equal(it, exp["iters"])
    assert np.array_equal(obj, exp["obj"])
    assert np.array_equal(cond, log10(exp["cond"]), equal_nan=True)
    ref = O.fit_pairs(codes, afk, nsel=20, order=O.ORDER_REF)
    lc = np.log10(ref["cond"])
    well = np.isfinite(lc) & (lc < 12)
    assert np.abs(jac - ref["jacquard"]).max(1)[well].max() <= 1e-6


def test_cli_samples_beyond_cohort(tmp_path):
    c = synth.generate(5, 200, seed=23, missing=0.0)
    vcf = str(tmp_path / "c.vcf")
    synth.write_vcf(c, vcf, "GT")
    r = subprocess.run([KESTREL, vcf, str(tmp_path / "o.tsv"), "-s", "6"], capture_output=True, text=True)
    assert r.returncode != 0 and "exceeds total samples" in r.stderr


@pytest.mark.parametrize("field", ["GL", "PL"])
def test_cli_soft_likelihoods_match_oracle(O, tmp_path, field):
    """Soft GL / PL rows take the likelihood path; the oracle fits the rows the
    CLI parsed (written back by --dump-probs)."""
    print(f"case: cli -f {field} soft", flush=True)
    c = synth.generate(24, 2500, seed=24, n_family_pairs=5, missing=0.01)
    gl = synth.soft_likelihoods(c.codes, seed=25, mean_depth=5.0)
    vcf = str(tmp_path / f"s.{field}.vcf.gz")
    synth.write_vcf(c, vcf, field, gl=gl)
    probs = str(tmp_path / "p.bin")
    r = subprocess.run([KESTREL, vcf, "/dev/null", "-f", field, "--dump-probs", probs], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = np.fromfile(probs, np.float64).reshape(gl.shape).transpose(1, 0, 2)  # (N, V, 3)
    rows, _ = run(vcf, str(tmp_path / "o.tsv"), "-f", field)
    x, y, kin, jac, it, obj, cond = as_arrays(rows, c.names)
    af = O.allele_freqs_gl(g)
    kept = O.keep_sites(af)
    gk = np.ascontiguousarray(g[:, kept])
    seg = np.minimum(segment_bounds(int(kept.sum())), gk.shape[1]).astype(np.uint32)
    exp = O.fit_pairs_gl(gk, af[kept], order=O.ORDER_SEG, seg=seg)
    assert np.array_equal(jac, exp["jacquard"]) and np.array_equal(it, exp["iters"])
    assert np.array_equal(obj, exp["obj"])
    assert np.array_equal(cond, log10(exp["cond"]), equal_nan=True)


@pytest.mark.parametrize("field", ["GL", "PL"])
def test_cli_multiallelic_matches_oracle(O, tmp_path, field):
    """2-, 3- and 4-allele records: the CLI buckets, re-buckets and fits them
    as the oracle does on the rows it parsed."""
    print(f"case: cli -f {field} multi-allelic", flush=True)
    b = synth.multi_likelihoods(16, (600, 150, 60), seed=61)
    vcf = str(tmp_path / f"m.{field}.vcf")
    synth.write_vcf_multi(vcf, b, field)
    probs = str(tmp_path / "p.bin")
    r = subprocess.run([KESTREL, vcf, "/dev/null", "-f", field, "--dump-probs", probs], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(probs, np.float64)
    shapes = [(600, 16, 3), (150, 16, 6), (60, 16, 10)]
    parsed, off = [], 0
    for s in shapes:
        k = int(np.prod(s))
        parsed.append(raw[off:off + k].reshape(s))
        off += k
    assert off == raw.size
    rows, log = run(vcf, str(tmp_path / "o.tsv"), "-f", field)
    assert "Found 150 tri-allelic sites" in log and "Found 60 tetra-allelic sites" in log
    x, y, kin, jac, it, obj, cond = as_arrays(rows, [f"S{i:05d}" for i in range(16)])
    parts, _ = O.multi_prepare({2: parsed[0], 3: parsed[1], 4: parsed[2]})
    pads, seg, base = {}, [0], 0
    for A in (2, 3, 4):
        nA = parts[A][0].shape[0]
        if A > 2 and nA == 0:
            continue
        bnd = segment_bounds(nA)
        pads[A] = int(bnd[-1])
        seg += [int(v) + base for v in bnd[1:]]
        base += int(bnd[-1])
    exp = O.multi_fit_pairs(parts, np.stack([x, y], 1), order=O.ORDER_SEG, pads=pads, seg=np.array(seg, np.uint32))
    assert np.array_equal(jac, exp["jacquard"]) and np.array_equal(it, exp["iters"])
    assert np.array_equal(obj, exp["obj"])
    assert np.array_equal(cond, log10(exp["cond"]), equal_nan=True)
    assert f"Identified {parts[3][0].shape[0]} tri-allelic sites" in log


def test_run_cli_from_python_writes_the_same_tsv(tmp_path):
    """kestrel_amd.run_cli (lib.rs:111-117) runs the same command line: the
    TSV is byte-identical to the binary's."""
    from kestrel_amd import run_cli
    print("case: run_cli from Python vs the binary", flush=True)
    c = synth.generate(24, 1500, seed=23, n_family_pairs=4, missing=0.01)
    vcf = str(tmp_path / "c.vcf.gz")
    synth.write_vcf(c, vcf, "GL")
    run(vcf, str(tmp_path / "a.tsv"), "-f", "GL", "-i", "50")
    assert run_cli(["kestrel", vcf, str(tmp_path / "b.tsv"), "-f", "GL", "-i", "50"]) is None
    assert open(tmp_path / "a.tsv", "rb").read() == open(tmp_path / "b.tsv", "rb").read()
