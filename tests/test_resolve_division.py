"""The ARNR resolve divides (16 * ctr + acc + ws / 2) by the weight sum ws
with a float reciprocal: floor(n / ws) = trunc(fma(n, rcp(ws), 2^-12)) for
ws in [16, 1024] (16 + 16 * (frames - 1), frames <= 64) and n < 256 * ws.
The device's v_rcp_f32 is within 1 ulp of 1 / ws; this checks every ws with
the reciprocal off by up to 2 ulp either way.  The computed quotient is
monotone in n, so both ends of every quotient's interval [q * ws, q * ws +
ws - 1] suffice."""
import numpy as np


def test_reciprocal_division_is_exact():
    bias = np.float32(2.0 ** -12)
    q = np.arange(256, dtype=np.int64)
    for d in range(16, 1025):
        n = np.concatenate([q * d, q * d + d - 1])
        want = n // d
        r = np.float32(1.0) / np.float32(d)
        rs = [r]
        for direction in (np.inf, -np.inf):
            x = r
            for _ in range(2):
                x = np.nextafter(x, np.float32(direction), dtype=np.float32)
                rs.append(x)
        for rr in rs:
            # fma: the product and the sum rounded once (exact in float64)
            got = (n.astype(np.float64) * np.float64(rr) + np.float64(bias)).astype(np.float32).astype(np.int64)
            assert np.array_equal(got, want), (d, rr)


def test_plain_reciprocal_needs_the_bias():
    # without the bias an exact quotient can come out one low: the check
    # above is not vacuous
    bias_free_fail = False
    for d in range(16, 1025):
        r = np.nextafter(np.float32(1.0) / np.float32(d), np.float32(0), dtype=np.float32)
        n = np.arange(256, dtype=np.int64) * d
        got = (n.astype(np.float64) * np.float64(r)).astype(np.float32).astype(np.int64)
        if (got != n // d).any():
            bias_free_fail = True
            break
    assert bias_free_fail
