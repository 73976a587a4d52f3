"""SplitMix64 against its published known-answer outputs, and the derived draws."""
import numpy as np

from lfric_core_amd.rng import Stream, splitmix64


def test_splitmix64_known_answers():
    # First outputs of SplitMix64 from seed 0 (Steele, Lea & Flood 2014; the
    # reference sequence used by xoshiro seeding).
    assert [int(v) for v in splitmix64(0, 3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert int(splitmix64(1234567, 1)[0]) == 6457827717110365317


def test_stream_is_sequential():
    a = Stream(7).raw(10)
    s = Stream(7)
    b = np.concatenate([s.raw(3), s.raw(7)])
    np.testing.assert_array_equal(a, b)


def test_uniform_and_normal_moments():
    s = Stream(3)
    u = s.uniform(200_000)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1.0 / 3.0) < 0.01
    z = Stream(4).normal(200_001)
    assert z.size == 200_001
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_torch_stream_matches_numpy_stream():
    import torch
    from lfric_core_amd.rng import TorchStream
    a, b = Stream(11), TorchStream(11, "cpu", chunk=333)
    assert np.array_equal(a.raw(1000), b.raw(1000).numpy().view(np.uint64))
    assert np.array_equal(a.uniform(501, -2.0, 3.0), b.uniform(501, -2.0, 3.0).numpy())
    np.testing.assert_allclose(b.normal(999).numpy(), a.normal(999), rtol=0, atol=1e-14)
    assert a.pos == b.pos
