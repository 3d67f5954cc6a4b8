"""Host-side sizing of the SAP digit output (fhetoothpaste_amd.operators.Context.sap_digit_counts):
the modulus rule of SAPDecomposer::Decompose (sap_decomposer.cpp:427-433) on hand-checked inputs and
against the oracle's own digit count."""
import numpy as np

from fhetoothpaste_amd.operators import Context


def test_digit_counts_closed_forms():
    n = 10
    rows = np.zeros((6, n + 1), dtype=np.uint64)
    rows[0, 0] = (1 << 40) - 1          # modulus 2^40, no value reaches it: 8 digits of 5 bits
    rows[1, 0] = 1 << 20                # 2^20 exactly: IntLog2 = 20, modulus 2^20 ...
    rows[1, 1] = 1 << 20                # ... which a_1 reaches: 2^21 -> 5 digits
    rows[2, 0] = (1 << 63) + 5          # 1 << IntLog2 = 1 << 64, taken as 1 (x86 shift), doubled by each a_i >= it
    rows[2, 1:] = 1 << 62               # 10 doublings: 2^10 -> 2 digits
    rows[3, :] = 0                      # modulus 1: no digit
    rows[4, 0] = 3                      # 4, then a_1 = 100 >= 4 -> 8, a_2 = 100 >= 8 -> 16: 1 digit
    rows[4, 1] = 100
    rows[4, 2] = 100
    rows[5, 0] = 1 << 33                # 2^33 -> 7 digits
    got = Context.sap_digit_counts(rows, 32)
    assert list(got) == [8, 5, 2, 0, 1, 7]
    assert list(Context.sap_digit_counts(rows, 2)) == [40, 21, 10, 0, 4, 33]


def test_digit_counts_match_oracle(oracle):
    N, n, radix = 1024, 4, 32
    lk = oracle.secret(n, 0, 7, 1)
    s = oracle.secret(N, 0, 7, 2)
    brk = oracle.brk(N, lk, s, n, 16, 4, oracle.BINARY, 0, 0.0, 7)
    keys = oracle.auto_keys(N, s, 16, 4, 0.0, 7)
    rng = np.random.default_rng(7)
    rows = [rng.integers(0, 1 << 12, size=n + 1, dtype=np.uint64),
            rng.integers(0, 1 << 7, size=n + 1, dtype=np.uint64)]
    rows[1][0] = 2
    rows = np.stack(rows)
    want = [oracle.sap_decompose(N, r, radix, 1, brk, n, 16, 4, oracle.BINARY, keys, 16, 4).shape[0] for r in rows]
    assert list(Context.sap_digit_counts(rows, radix)) == want
