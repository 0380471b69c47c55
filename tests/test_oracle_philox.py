"""The oracle's Philox4x32-10 weight stream (weights of the differentiation workload's
L = sum w K3^2) on the CPU.

Pinned by the generator's published known-answer vectors (Random123 distribution,
kat_vectors: philox4x32 with 10 rounds) and by the element -> (counter, word) rule that
the device kernel follows."""
import numpy as np

from oracle import fc3_oracle as O


def _u32(*a):
    return np.array(a, dtype=np.uint32)


def test_philox4x32_10_known_answers():
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = O.philox4x32_10(_u32(*ctr)[None, :], _u32(*key)[None, :])[0]
        assert [int(x) for x in got] == list(want)


def test_weight_stream_indexing():
    seed = (7 << 32) | 12345
    w = O.philox_uniform(seed, 40)
    for e in (0, 1, 5, 39):
        q = e // 4
        words = O.philox4x32_10(_u32(q & 0xFFFFFFFF, q >> 32, 0, 0)[None, :],
                                _u32(seed & 0xFFFFFFFF, seed >> 32)[None, :])[0]
        assert w[e] == np.float32((int(words[e % 4]) >> 8) * 2.0 ** -24)
    # any window of the stream equals the same slice of a longer one
    assert np.array_equal(O.philox_uniform(seed, 9, start=31), w[31:40])


def test_weights_uniform_and_loss_definition():
    w = O.philox_uniform(99, 200000)
    assert w.dtype == np.float32 and 0.0 <= w.min() and w.max() < 1.0
    assert abs(float(w.mean()) - 0.5) < 3e-3 and abs(float(w.var()) - 1 / 12) < 2e-3
    k3 = np.asfortranarray(np.random.default_rng(3).standard_normal((4, 5, 3)).astype(np.float32))
    L, g = O.weighted_square_loss(k3, 99)
    wf = O.philox_uniform(99, 60).reshape((4, 5, 3), order="F").astype(np.float64)
    assert np.isclose(L, float(np.sum(wf * k3.astype(np.float64) ** 2)), rtol=1e-14)
    assert np.array_equal(g, (np.float32(2) * wf.astype(np.float32)) * k3)
