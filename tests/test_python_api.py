"""Python array API (lib.rs:26-161): input conversion and array assembly,
checked on the CPU without a device."""
import numpy as np
import pytest

from kestrel_amd import OUTPUT_DTYPE, build_arrays, relatedness_coefficients_gl
from kestrel_amd import relatedness_coefficients_gt
from kestrel_amd.api import _gt_indices


def test_gt_indices_unphased_index():
    # coefficients.rs:27-32: j(j+1)/2 + i with i = min, j = max
    g = np.array([[[0, 0], [0, 1], [1, 0], [1, 1]]], np.uint8)  # one site, four samples
    assert _gt_indices(g).tolist() == [[0], [1], [1], [2]]
    g = np.array([[[0, 2], [2, 1], [2, 2], [3, 0], [1, 3], [3, 3]]], np.uint8)
    assert _gt_indices(g).ravel().tolist() == [3, 4, 5, 6, 7, 9]


def test_gt_indices_rejects_fifth_allele_and_bad_shape():
    with pytest.raises(ValueError):
        _gt_indices(np.array([[[0, 4]]], np.uint8))
    with pytest.raises(ValueError):
        _gt_indices(np.zeros((3, 4), np.uint8))


def test_gt_allele_counts_checked_before_any_device_call():
    g = np.zeros((2, 3, 2), np.uint8)
    g[1, 0] = (0, 2)
    with pytest.raises(ValueError, match="exceeds"):
        relatedness_coefficients_gt(g, alleles=[2, 2])
    with pytest.raises(ValueError, match="2, 3 or 4"):
        relatedness_coefficients_gt(g, alleles=[2, 5])


def test_gl_input_checks_before_any_device_call():
    with pytest.raises(NotImplementedError):
        relatedness_coefficients_gl(np.eye(3)[None].repeat(2, 0), allele_frequencies=np.full((2, 2), 0.5))
    with pytest.raises(ValueError, match="3, 6 or 10 genotypes"):
        relatedness_coefficients_gl(np.zeros((2, 3, 5)))


def test_build_arrays_matches_reference_assembly():
    # lib.rs:120-161: swap delta3<->delta5, delta4<->delta6 for (y, x), kinship
    # symmetric, inbreeding from the diagonal, zeros elsewhere
    out = np.zeros(3, OUTPUT_DTYPE)
    d = np.arange(1, 10) / 45.0
    out[0]["x"], out[0]["y"], out[0]["jacquard"] = 0, 1, d
    out[1]["x"], out[1]["y"], out[1]["jacquard"] = 1, 1, np.eye(9)[0]
    out[2]["x"], out[2]["y"], out[2]["jacquard"] = 0, 0, np.eye(9)[8]
    jac, kin, inb = build_arrays(out, 3)
    assert np.array_equal(jac[0, 1], d)
    assert np.array_equal(jac[1, 0], d[[0, 1, 4, 5, 2, 3, 6, 7, 8]])
    k01 = d @ np.array([1.0, 0.0, 0.5, 0.0, 0.5, 0.0, 0.5, 0.25, 0.0])
    assert kin[0, 1] == kin[1, 0] and abs(kin[0, 1] - k01) < 1e-15
    assert inb[1] == 1.0 and inb[0] == -1.0 and inb[2] == 0.0
    assert (jac[2] == 0).all() and (kin[2] == 0).all()


def test_run_cli_exits_and_errors_like_the_reference(tmp_path, capfd):
    """lib.rs:111-117 run_cli: clap's parse_from ends the process on a usage
    error (2) and after --version (0); a failed run raises RuntimeError with
    the error message.  None of these reaches the device."""
    from kestrel_amd import run_cli
    with pytest.raises(SystemExit) as e:
        run_cli(["kestrel", "--version"])
    assert e.value.code == 0
    assert "kestrel" in capfd.readouterr().out
    with pytest.raises(SystemExit) as e:
        run_cli(["kestrel", str(tmp_path / "in.vcf"), str(tmp_path / "o.tsv"), "-f", "XX"])
    assert e.value.code == 2
    assert "invalid value" in capfd.readouterr().err
    with pytest.raises(RuntimeError, match="No such file"):
        run_cli(["kestrel", str(tmp_path / "missing.vcf"), str(tmp_path / "o.tsv")])
