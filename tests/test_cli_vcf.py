"""CPU tests of the kestrel CLI's VCF decoding (vcf.rs:96-482 rules) and of
its TSV float text (ryu, the csv crate's f64 writer).  Decoding is compared
bit-exactly: codes 0/1/2 = unphased genotype index (coefficients.rs:18-38),
3 = missing."""
import gzip
import math
import os
import subprocess

import numpy as np
import pytest

from kestrel_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KESTREL = os.path.join(ROOT, "kestrel_amd", "kestrel")

HEADER = ("##fileformat=VCFv4.2\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\tA\tB\tC\n")


def dump(path, field="GT", af_info=False, expect_ok=True):
    cmd = [KESTREL, path, "/dev/null", "-f", field, "--dump-codes"] + (["-a", "info"] if af_info else [])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if not expect_ok:
        assert r.returncode != 0
        return r.stderr
    assert r.returncode == 0, r.stderr
    lines = r.stdout.rstrip("\n").split("\n")
    samples = lines[0].split("\t")
    sites = [ln.split("\t") for ln in lines[1:]]
    counts = dict(kv.split("=") for kv in r.stderr.strip().split("\n")[-1].split())
    return samples, sites, {k: int(v) for k, v in counts.items()}


def write(tmp_path, body, name="t.vcf"):
    p = tmp_path / name
    p.write_text(HEADER + body)
    return str(p)


def test_gt_decoding_rules(tmp_path):
    body = (
        "1\t1\t.\tA\tG\t.\tPASS\t.\tGT\t0/0\t0|1\t1/1\n"       # kept
        "1\t2\t.\tA\tG\t.\tPASS\t.\tGT:DP\t1/0:5\t1|1:3\t0/0:1\n"  # unphased index, extra FORMAT key
        "1\t3\t.\tAT\tG\t.\tPASS\t.\tGT\t0/0\t0/1\t1/1\n"      # REF not a single base: not a SNP
        "1\t4\t.\tA\tGC\t.\tPASS\t.\tGT\t0/0\t0/1\t1/1\n"      # ALT not a single base
        "1\t5\t.\tA\t.\t.\tPASS\t.\tGT\t0/0\t0/0\t0/0\n"       # no ALT: one allele
        "1\t6\t.\ta\tt\t.\tPASS\t.\tGT\t./.\t0/1\t1/1\n"       # missing call drops the site (vcf.rs:217-221)
        "1\t7\t.\tC\tT\t.\tPASS\t.\tGT\t0/1\t.|1\t0/0\n"       # half-missing too
        "1\t8\t.\tc\tg\t.\tPASS\t.\tGT\t1/1\t0/1\t0/1\n"       # lower-case bases are SNPs
    )
    samples, sites, counts = dump(write(tmp_path, body))
    assert samples == ["A", "B", "C"]
    assert [s[0] for s in sites] == ["012", "120", "211"]
    assert counts == {"total": 8, "not_snp": 3, "skipped_missing": 2}


def test_gt_errors(tmp_path):
    err = dump(write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGT\t0\t0/1\t1/1\n"), expect_ok=False)
    assert "not diploid" in err
    err = dump(write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGT\t0/1/1\t0/1\t1/1\n"), expect_ok=False)
    assert "not diploid" in err
    err = dump(write(tmp_path, "1\t1\t.\tA\tG,T\t.\tPASS\t.\tGT\t0/3\t0/1\t1/1\n"), expect_ok=False)
    assert "allele index 3 exceeds" in err
    err = dump(write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGT\t0/2\t0/1\t1/1\n"), expect_ok=False)
    assert "allele index 2 exceeds" in err
    err = dump(write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tDP\t1\t2\t3\n"), expect_ok=False)
    assert "No GT data found" in err


def test_gl_pl_decoding(tmp_path):
    body = (
        "1\t1\t.\tA\tG\t.\tPASS\tAF=0.25\tGL:PL\t0,-1000,-1000:0,9999,9999\t-1000,0,-1000:9999,0,9999\t"
        "-2000,-2000,-1000:9999,9999,0\n"
        "1\t2\t.\tA\tG\t.\tPASS\tAF=0.5\tGL:PL\t.,.,.:.,.,.\t-1000,0,-1000:9999,0,9999\t0,-1000,.:0,9999,.\n"
    )
    p = write(tmp_path, body)
    for field in ("GL", "PL"):
        _, sites, _ = dump(p, field, af_info=True)
        # a NaN in any likelihood drops the row (vcf.rs:355,397; lanevector.rs:100-104)
        assert [s[0] for s in sites] == ["012", "313"]
        assert sites[0][1:] == ["0.75", "0.25"] and sites[1][1:] == ["0.5", "0.5"]
    # soft rows are kept as likelihoods: PL -> 10^(-pl/10) (vcf.rs:407), GL ->
    # 10^(gl - max gl) over the float32 values (vcf.rs:349-368); libm pow
    soft = write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGL:PL\t-0.3,0,-2.7:0,30,300\t"
                           "-1.25,-0.5,.:3,0,7\t0,-1000,-1000:9999,0,9999\n", "s.vcf")
    f32 = lambda t: float(np.float32(t))
    exp = {"PL": [[10 ** -0.0, math.pow(10.0, -3.0), math.pow(10.0, -30.0)],
                  [math.pow(10.0, -0.3), 1.0, math.pow(10.0, -0.7)],
                  [math.pow(10.0, -999.9), 1.0, math.pow(10.0, -999.9)]],
           "GL": [[math.pow(10.0, f32("-0.3") - 0.0), 1.0, math.pow(10.0, f32("-2.7") - 0.0)],
                  [math.pow(10.0, f32("-1.25") - f32("-0.5")), 1.0, math.nan],
                  [1.0, 0.0, 0.0]]}
    for field in ("GL", "PL"):
        out = str(tmp_path / f"p.{field}.bin")
        r = subprocess.run([KESTREL, soft, "/dev/null", "-f", field, "--dump-probs", out], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(out, np.float64).reshape(1, 3, 3)[0]
        assert np.array_equal(got, np.array(exp[field]), equal_nan=True), (field, got)
    bad = write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGL\t0,-1000\t0,-1,-1\t0,-1,-1\n", "b.vcf")
    assert "Not enough genotype likelihoods" in dump(bad, "GL", expect_ok=False)


def test_synthetic_vcf_roundtrip_gz(tmp_path):
    c = synth.generate(17, 300, seed=9, missing=0.05)
    for field in ("GT", "GL", "PL"):
        path = str(tmp_path / f"c.{field}.vcf.gz")
        synth.write_vcf(c, path, field)
        samples, sites, counts = dump(path, field)
        assert samples == c.names
        got = np.array([[int(ch) for ch in s[0]] for s in sites], np.uint8).T
        if field == "GT":
            keep = (c.codes != 3).all(0)
            assert np.array_equal(got, c.codes[:, keep])
            assert counts["skipped_missing"] == int((~keep).sum())
        else:
            assert np.array_equal(got, c.codes)


def test_float_text_matches_ryu():
    vals = ["0.25", "1e-8", "1.5e-7", "100", "1e15", "1e16", "0.00001", "0.000001", "123456789012345680",
            "-0.5", "nan", "inf", "-inf", "0.1", "0.30000000000000004", "5e-324", "1.7976931348623157e308"]
    exp = ["0.25", "1e-8", "1.5e-7", "100.0", "1000000000000000.0", "1e16", "0.00001", "1e-6",
           "1.2345678901234568e17", "-0.5", "NaN", "inf", "-inf", "0.1", "0.30000000000000004", "5e-324",
           "1.7976931348623157e308"]
    r = subprocess.run([KESTREL, "--format-floats"], input="\n".join(vals) + "\n", capture_output=True, text=True,
                       check=True)
    assert r.stdout.split("\n")[:-1] == exp


def test_usage_errors():
    r = subprocess.run([KESTREL, "only_one_arg"], capture_output=True, text=True)
    assert r.returncode == 2 and "required" in r.stderr
    r = subprocess.run([KESTREL, "a", "b", "-f", "XX"], capture_output=True, text=True)
    assert r.returncode == 2 and "invalid value" in r.stderr
    r = subprocess.run([KESTREL, "/nonexistent.vcf", "out.tsv"], capture_output=True, text=True)
    assert r.returncode == 1 and "No such file" in r.stderr


def _dump_all(path, field, threads):
    r = subprocess.run([KESTREL, path, "/dev/null", "-f", field, "--dump-codes", "-t", str(threads)],
                       capture_output=True, text=True)
    return r.returncode, r.stdout, r.stderr.strip().split("\n")[-1]


@pytest.mark.parametrize("field", ["GT", "PL"])
def test_parallel_reader_matches_serial(tmp_path, field):
    """The same records for 1 or 8 threads and plain, gzip or BGZF input."""
    c = synth.generate(100, 30000, seed=41, missing=0.00002)
    files = []
    for name, kw in (("a.vcf", {}), ("b.vcf.gz", {}), ("c.vcf.gz", {"bgzf": True})):
        p = str(tmp_path / name)
        synth.write_vcf(c, p, field, **kw)
        files.append(p)
    base = _dump_all(files[0], field, 1)
    assert base[0] == 0 and base[1].count("\n") > 1000
    for p in files:
        for t in (1, 3, 8):
            assert _dump_all(p, field, t) == base, (p, t)


def test_parallel_reader_reports_first_error(tmp_path):
    """An error deep in the file is found and numbered as a serial read would."""
    c = synth.generate(30, 20000, seed=42, missing=0.0)
    p = str(tmp_path / "e.vcf")
    synth.write_vcf(c, p, "GT")
    lines = open(p).read().split("\n")
    hdr = sum(1 for ln in lines if ln.startswith("#"))
    lines[hdr + 15000] = lines[hdr + 15000].replace("0/1", "0/x", 1) if "0/1" in lines[hdr + 15000] else \
        lines[hdr + 15000].replace("0/0", "0/x", 1)
    lines[hdr + 17000] = "\t".join(lines[hdr + 17000].split("\t")[:5])  # a later, different error
    open(p, "w").write("\n".join(lines))
    outs = {t: _dump_all(p, "GT", t) for t in (1, 8)}
    assert outs[1] == outs[8]
    assert outs[1][0] != 0
    r = subprocess.run([KESTREL, p, "/dev/null", "--dump-codes", "-t", "8"], capture_output=True, text=True)
    assert "Error reading genotype" in r.stderr


def test_last_line_without_newline(tmp_path):
    p = write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\t.\tGT\t0/0\t0/1\t1/1", "n.vcf")
    _, sites, counts = dump(p)
    assert [s[0] for s in sites] == ["012"] and counts["total"] == 1


def test_multiallelic_likelihoods_are_bucketed(tmp_path):
    """GL / PL records with 3 and 4 alleles keep their 6 / 10 likelihoods in
    their own buckets (vcf.rs:270-305), after the biallelic rows."""
    body = ("1\t1\t.\tA\tG\t.\tPASS\t.\tPL\t0,10,20\t5,0,5\t20,10,0\n"
            "1\t2\t.\tA\tG,T\t.\tPASS\t.\tPL\t0,1,2,3,4,5\t9,8,7,6,5,0\t.,.,.,.,.,.\n"
            "1\t3\t.\tA\tC,G,T\t.\tPASS\t.\tPL\t" + ",".join(str(i) for i in range(10)) + "\t"
            + ",".join(str(20 - i) for i in range(10)) + "\t0,0,0,0,0,0,0,0,0,0\n")
    p = write(tmp_path, body, "m.vcf")
    out = str(tmp_path / "p.bin")
    r = subprocess.run([KESTREL, p, "/dev/null", "-f", "PL", "--dump-probs", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float64)
    assert got.size == 3 * 3 + 3 * 6 + 3 * 10
    pl2 = [[0, 10, 20], [5, 0, 5], [20, 10, 0]]
    pl3 = [[0, 1, 2, 3, 4, 5], [9, 8, 7, 6, 5, 0]]
    pl4 = [list(range(10)), [20 - i for i in range(10)], [0] * 10]
    exp = [math.pow(10.0, -v / 10.0) for row in pl2 for v in row]
    exp += [math.pow(10.0, -v / 10.0) for row in pl3 for v in row] + [math.nan] * 6
    exp += [math.pow(10.0, -v / 10.0) for row in pl4 for v in row]
    assert np.array_equal(got, np.array(exp), equal_nan=True)
    bad = write(tmp_path, "1\t1\t.\tA\tG,T\t.\tPASS\t.\tGL\t0,0,0\t0,0,0,0,0,0\t0,0,0,0,0,0\n", "b.vcf")
    assert "Not enough genotype likelihoods" in dump(bad, "GL", expect_ok=False)


def dump_multi(path, af_info=False):
    """--dump-codes of a GT file: (2-allele lines, {3: lines, 4: lines})."""
    r = subprocess.run([KESTREL, path, "/dev/null", "-f", "GT", "--dump-codes"] + (["-a", "info"] if af_info else []),
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.rstrip("\n").split("\n")[1:]
    two = [ln.split("\t") for ln in lines if ":" not in ln.split("\t")[0]]
    multi = {3: [], 4: []}
    for ln in lines:
        head = ln.split("\t")[0]
        if ":" in head:
            A, codes = head.split(":")
            multi[int(A)].append([codes] + ln.split("\t")[1:])
    return two, multi


def test_gt_multiallelic_sites_are_bucketed(tmp_path):
    """GT records with 3 and 4 alleles go to their own buckets (vcf.rs:126-160)
    as genotype indices j(j+1)/2 + i (coefficients.rs:19-38); a missing call
    drops the site in every bucket; INFO AF holds A - 1 values and the
    reference allele gets 1 - clamp(sum) (vcf.rs:446-482)."""
    body = ("1\t1\t.\tA\tG\t.\tPASS\tAF=0.25\tGT\t0/0\t0/1\t1/1\n"
            "1\t2\t.\tA\tG,T\t.\tPASS\tAF=0.25,0.5\tGT\t0/2\t2|1\t2/2\n"
            "1\t3\t.\tA\tC,G,T\t.\tPASS\tAF=0.1,0.2,0.3\tGT\t3/0\t1/3\t3/3\n"
            "1\t4\t.\tA\tG,T\t.\tPASS\tAF=0.5,0.5\tGT\t0/0\t./2\t1/1\n"
            "1\t5\t.\tA\tG,T\t.\tPASS\tAF=0.75,0.5\tGT\t1/0\t1/1\t0/0\n")
    p = write(tmp_path, body, "m.vcf")
    two, multi = dump_multi(p, af_info=True)
    assert [t[0] for t in two] == ["012"]
    assert [m[0] for m in multi[3]] == ["345", "120"]
    assert [m[0] for m in multi[4]] == ["679"]
    f32 = lambda x: float(np.float32(x))
    assert [float(v) for v in multi[3][0][1:]] == [1.0 - (f32(0.25) + f32(0.5)), f32(0.25), f32(0.5)]
    assert [float(v) for v in multi[3][1][1:]] == [0.0, f32(0.75), f32(0.5)]  # clamped alt sum
    assert [float(v) for v in multi[4][0][1:]] == [1.0 - (f32(0.1) + f32(0.2) + f32(0.3)), f32(0.1), f32(0.2),
                                                   f32(0.3)]
    _, counts = dump(p)[1:]
    assert counts == {"total": 5, "not_snp": 0, "skipped_missing": 1}
    for af, msg in (("AF=0.1", "Not enough allele frequencies"), ("AF=0.1,0.2,0.3", "Too many allele frequencies")):
        bad = write(tmp_path, f"1\t1\t.\tA\tG,T\t.\tPASS\t{af}\tGT\t0/2\t0/1\t1/1\n", "b.vcf")
        assert msg in dump(bad, "GT", af_info=True, expect_ok=False)


def test_gt_multiallelic_synthetic_roundtrip(tmp_path):
    idx = synth.multi_genotypes(9, (25, 12, 7), seed=4, missing=0.02)
    p = str(tmp_path / "g.vcf.gz")
    synth.write_vcf_gt_multi(p, idx)
    two, multi = dump_multi(p)
    for A, got, b in ((2, [t[0] for t in two], idx[0]), (3, [m[0] for m in multi[3]], idx[1]),
                      (4, [m[0] for m in multi[4]], idx[2])):
        keep = ~(b == 255).any(axis=0)
        exp = ["".join(str(v) for v in b[:, s]) for s in np.flatnonzero(keep)]
        assert got == exp, A


def test_float_fields_parse_as_f32(tmp_path):
    """INFO AF (and GL) values are f32 text (noodles' Float): a leading '+' is
    accepted, values round correctly to f32, below f32's range give 0 and
    above it inf; a malformed value is an error."""
    body = ("1\t1\t.\tA\tG\t.\tPASS\tAF=+0.25\tGT\t0/0\t0/1\t1/1\n"
            "1\t2\t.\tA\tG\t.\tPASS\tAF=1e-50\tGT\t0/0\t0/1\t1/1\n"
            "1\t3\t.\tA\tG\t.\tPASS\tAF=1e39\tGT\t0/0\t0/1\t1/1\n"
            "1\t4\t.\tA\tG\t.\tPASS\tAF=0.1\tGT\t0/0\t0/1\t1/1\n")
    _, sites, _ = dump(write(tmp_path, body), af_info=True)
    alt = [float(s[2]) for s in sites]
    assert alt[0] == 0.25 and alt[1] == 0.0 and math.isinf(alt[2]) and alt[3] == float(np.float32(0.1))
    assert [float(s[1]) for s in sites] == [0.75, 1.0, 0.0, 1.0 - float(np.float32(0.1))]
    err = dump(write(tmp_path, "1\t1\t.\tA\tG\t.\tPASS\tAF=0.1x\tGT\t0/0\t0/1\t1/1\n", "b.vcf"), af_info=True,
               expect_ok=False)
    assert "Error reading AF" in err


def _bgzf_block(payload: bytes, isize=None, bsize=None) -> bytes:
    import struct
    import zlib
    co = zlib.compressobj(6, zlib.DEFLATED, -15)
    comp = co.compress(payload) + co.flush()
    bs = 18 + len(comp) + 8 - 1 if bsize is None else bsize
    head = b"\x1f\x8b\x08\x04" + b"\x00" * 4 + b"\x00\xff" + struct.pack("<H", 6) + b"BC" + struct.pack("<HH", 2, bs)
    return head + comp + struct.pack("<II", zlib.crc32(payload) & 0xFFFFFFFF,
                                     len(payload) if isize is None else isize)


def test_bgzf_corrupt_block_sizes(tmp_path):
    """A BGZF block whose BSIZE is shorter than its own header and footer, or
    whose ISIZE exceeds the 64 KiB a block may hold, is an error with a
    message (not a read past the block or a huge allocation)."""
    body = (HEADER + "1\t1\t.\tA\tC\t.\t.\t.\tGT\t0/1\t1/1\t0/0\n").encode()
    eof = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
    good = tmp_path / "good.vcf.gz"
    good.write_bytes(_bgzf_block(body) + eof)
    samples, sites, _ = dump(str(good))
    assert samples == ["A", "B", "C"] and len(sites) == 1
    for name, blk in (("short.vcf.gz", _bgzf_block(body, bsize=5)),
                      ("huge.vcf.gz", _bgzf_block(body, isize=1 << 31))):
        p = tmp_path / name
        p.write_bytes(blk + eof)
        err = dump(str(p), expect_ok=False)
        assert "bgzf" in err, err
