"""The record rules of the exact sums by chunk programs, on the CPU.

tests/walk_model.py restates the rules the device programs follow
(hera_amd/csrc/dev/walk_sum.hip: runs on a binade's or a row's unit, first
ties added by themselves, start intervals and grid checks) and the checked
in-order walk.  These tests hold the method itself to the sequential sum,
bit for bit, on adversarial rows -- also when every chunk's guessed start
is wrong, which only moves work to the value-by-value path.  The device
kernels are held to the same sums by tests/test_gpu_sums.py.
"""
import numpy as np
import pytest

from walk_model import exact_sum


def seq_sum(x):
    return np.cumsum(np.concatenate([[0.0], x]))[-1]


def same_bits(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return np.float64(a).view(np.int64) == np.float64(b).view(np.int64)


def rows():
    rng = np.random.default_rng(11)
    out = {}
    out["normal"] = rng.normal(size=1500)
    out["lognormal_wide"] = rng.lognormal(0, 12, size=1500)
    big = rng.lognormal(0, 3, size=1000)
    out["return_to_zero"] = np.concatenate([big, -rng.permutation(big)]) + \
        rng.normal(scale=1e-9, size=2000)
    ties = rng.integers(0, 4, size=1200) + 0.5
    ties[0] = 2.0 ** 52
    ties[1::2] *= -1.0
    out["ties"] = ties
    hov = rng.integers(-3, 4, size=1500) * 2.0 ** -53
    hov[0] = 1.0
    out["boundary_hover"] = hov
    out["short_walk"] = rng.integers(-2000, 2001, size=3000) * 2.0 ** -60
    mixed = rng.integers(-64, 65, size=3000) * 2.0 ** -40
    mixed[::5] = rng.normal(size=600) * 2.0 ** -34
    out["short_mixed"] = mixed
    z = np.zeros(600)
    z[::3] = -0.0
    out["signed_zeros"] = z
    out["subnormal"] = rng.normal(size=1000) * 1e-310
    out["huge"] = rng.normal(size=800) * 2.0 ** 1010
    inf = rng.normal(size=700)
    inf[300] = np.inf
    out["nonfinite"] = inf
    return out


ROWS = rows()


@pytest.mark.parametrize("grid", [True, False])
@pytest.mark.parametrize("name", sorted(ROWS))
def test_model_bitwise(name, grid):
    x = ROWS[name]
    got, nrec, fails = exact_sum(x, L=256, row=16, use_grid=grid)
    assert same_bits(got, seq_sum(x)), (got, seq_sum(x))
    # and the squares, as the sums of squares run
    got2, _, _ = exact_sum(x * x, L=256, row=16, use_grid=False)
    assert same_bits(got2, seq_sum(x * x))


@pytest.mark.parametrize("name", ["normal", "return_to_zero", "short_walk",
                                  "short_mixed", "ties", "boundary_hover"])
def test_model_exact_with_wrong_guesses(name):
    """Guesses off by a few units of the running sums, or by far more:
    the records are built for the wrong trajectory, the walk's checks
    catch the segments whose true starts they do not fit, and the sum
    stays the sequential one."""
    x = ROWS[name]
    scale = float(np.max(np.abs(np.cumsum(x)))) or 1.0
    total_fails = 0
    for shift in (1e-12, 1e-6, 0.3):
        err = lambda j: ((-1) ** j) * shift * scale * (j + 1)
        got, nrec, fails = exact_sum(x, L=256, row=16, guess_error=err)
        total_fails += fails
        assert same_bits(got, seq_sum(x)), (shift, got, seq_sum(x))
    # the checks were exercised, not only passed
    assert total_fails > 0


def test_row_grid_cuts_records_near_zero():
    """Short values whose running sum crosses binades around zero: runs on
    a row's grid span the crossings that cut binade runs."""
    x = ROWS["short_walk"]
    _, rec_grid, _ = exact_sum(x, L=256, row=16, use_grid=True)
    _, rec_binade, _ = exact_sum(x, L=256, row=16, use_grid=False)
    assert rec_grid < 0.8 * rec_binade, (rec_grid, rec_binade)


U = 2.0 ** -52


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_interval_stops_a_run_at_the_binade_top(sign):
    """A guess 2^-20 below the truth makes six steps of 1.25 ulps look like
    one run inside [1, 2); truly they cross 2.0, where the ulp doubles and
    the rounding changes.  The segment's start interval (partial sums at
    most 2^53 - 1 units) rejects the true start, and the values are added
    one by one."""
    x = sign * np.array([2 - 3 * U] + [1.25 * U] * 6)
    got, nrec, fails = exact_sum(x, L=256, row=16,
                                 guess_error=lambda j: -sign * 2.0 ** -20)
    assert same_bits(got, seq_sum(x))
    assert fails == 1


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_grid_check_rejects_an_off_grid_start(sign):
    """A guess 2^-20 above the truth makes 16 - 8 ulps look like a start in
    [16, 32), an anchor for a run on that binade's unit 16 ulps; the true
    start is not a multiple of it, and adding 48 ulps at once would round
    to another neighbour than adding 32 and 16 in turn (16 + 32 ulps at
    once, 16 + 48 in turn).  The start's grid check sends the run value by
    value."""
    x = sign * np.array([16 - 8 * U, 32 * U, 16 * U])
    got, nrec, fails = exact_sum(x, L=256, row=16,
                                 guess_error=lambda j: sign * 2.0 ** -20)
    assert same_bits(got, seq_sum(x))
    assert fails == 1
