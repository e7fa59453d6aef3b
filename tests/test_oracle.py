"""Pins the oracle (oracle/ddo.c) before anything is checked against it."""
import numpy as np
import pytest
from conftest import golden

BITS = np.array([1 << i for i in range(64)], dtype=np.uint64)


def popcounts(h):
    return np.array([[bin(int(x)).count("1") for x in row] for row in h])


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_oracle_equals_unpruned_minimax(oracle, k):
    """Exact search with memo and equivalence pruning == textbook minimax, all 20 results."""
    h = oracle.deals(31 + k, 40, k)
    assert np.array_equal(oracle.tables(h), oracle.tables(h, brute=True))


def test_oracle_equals_minimax_five_cards(oracle):
    h = oracle.deals(99, 6, 5)
    assert np.array_equal(oracle.tables(h), oracle.tables(h, brute=True))


def test_oracle_catches_a_planted_rule_error(oracle):
    """A table with one entry off by one must differ from the oracle's (the check has teeth)."""
    h = oracle.deals(5, 10, 4)
    t = oracle.tables(h)
    bad = t.copy()
    bad[3, 7] = (bad[3, 7] + 1) % 14
    assert not np.array_equal(bad, oracle.tables(h, brute=True))


def test_oracle_table_conventions(oracle):
    """Declarer d's leader is d+1; E/W entries are 13 - NS tricks even for endings."""
    h = oracle.deals(8, 30, 3)
    t = oracle.tables(h).astype(int).reshape(-1, 5, 4)
    for i in range(len(h)):
        for st in range(5):
            trump = -1 if st == 4 else 3 - st
            for d in range(4):
                ns = oracle.lib.ddo_ns_tricks(h[i].ctypes.data, trump, (d + 1) & 3)
                assert t[i, st, d] == (ns if d % 2 == 0 else 13 - ns)
    assert (t[:, :, 1::2] >= 10).all()  # 3-card endings: 13 - NS >= 10


def test_oracle_matches_reference_on_c1_endings(oracle):
    """1,024 of the 4,096 8-card endings vs the reference CPU solver's tables."""
    g = golden("c1_endings.npz")
    h = g["hands"][:1024]
    assert np.array_equal(oracle.tables(h), g["tables"][:1024])


def test_oracle_generator_is_the_product_generator(oracle, ddx):
    for seed, k in [(1, 8), (2026, 13), (7, 5)]:
        assert np.array_equal(oracle.deals(seed, 50, k, first=123), ddx.random_deals(seed, 50, k, 123))
    g = golden("c1_endings.npz")
    assert np.array_equal(oracle.deals(int(g["seed"]), 64, int(g["k"])), g["hands"][:64])
    g = golden("full_2000.npz")
    assert np.array_equal(oracle.deals(int(g["seed"]), 64, 13), g["hands"][:64])


def test_generator_deals_are_valid_and_uniform(oracle):
    h = oracle.deals(424242, 20000, 13)
    assert (popcounts(h[:200]) == 13).all()
    assert (np.bitwise_or.reduce(h, axis=1) == np.uint64(0x1FFF1FFF1FFF1FFF)).all()
    assert (h[:, 0] & h[:, 1] == 0).all() and (h[:, 2] & h[:, 3] == 0).all()
    # Each card lands in each seat with probability 1/4: chi-square over 52 cards x 4 seats.
    counts = np.zeros((64, 4))
    for s in range(4):
        counts[:, s] = ((h[:, s, None] & BITS[None, :]) != 0).sum(axis=0)
    c = counts[[16 * s + r for s in range(4) for r in range(13)]]
    chi2 = ((c - 5000.0) ** 2 / 5000.0).sum()
    assert chi2 < 230  # 156 d.o.f.: mean 156, sd 17.7


def test_pgx_decoding_matches_dds_fixture(oracle, ddx, tmp_path):
    a = golden("dds_4000.npy")
    hands, tables = oracle.pgx(a)
    assert (popcounts(hands[:500]) == 13).all()
    assert (np.bitwise_or.reduce(hands, axis=1) == np.uint64(0x1FFF1FFF1FFF1FFF)).all()
    assert tables.max() <= 13
    # The product's reader decodes the same file identically; its writer round-trips.
    p = str(tmp_path / "dds.npy")
    np.save(p, a)
    h2, t2 = ddx.pgx_read(p)
    assert np.array_equal(h2, hands) and np.array_equal(t2, tables)
    q = str(tmp_path / "rt.npy")
    ddx.pgx_write(q, hands[:100], tables[:100])
    assert np.array_equal(np.load(q), a[:, :100, :])


def test_mixed_fixture_deals_and_small_endings(oracle, ddx):
    """The mixed-endings fixture: its deals are the seeded generator's, and the oracle agrees with
    the reference solver's tables on its 5- and 7-card endings."""
    m = golden("mixed_endings.npz")
    for k in (5, 7, 9, 11):
        rows = m["k"] == k
        assert np.array_equal(m["hands"][rows], ddx.random_deals(3000 + k, int(rows.sum()), k))
    small = np.isin(m["k"], (5, 7))
    assert np.array_equal(oracle.tables(m["hands"][small]), m["tables"][small])
