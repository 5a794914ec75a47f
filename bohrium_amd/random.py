"""Counter-based random numbers (Random123 philox2x32-10) generated on the engine.

Mirrors the reference bh107 RandomState (bridge/bh107/bh107/random.py:40-190):
the key is the 32-bit seed, every draw consumes `index` counters, and a uniform
double is float64(u64) / float64(2**64 - 1).
"""
import numpy as np

from . import _bh_api as api
from .bharray import BhArray


class RandomState:
    def __init__(self, seed=None):
        self.seed(seed)

    def seed(self, seed=None):
        if seed is None:
            seed = np.random.SeedSequence().entropy
        self.key = int(np.uint32(int(seed) & 0xFFFFFFFF))
        self.index = 0

    def get_state(self):
        return ("Random123", self.index, self.key)

    def set_state(self, state):
        if len(state) != 3 or state[0] != "Random123":
            raise ValueError("state must be ('Random123', index, key)")
        self.index, self.key = int(state[1]), int(state[2])

    def random123(self, shape):
        """uint64 array of raw philox2x32-10 outputs (counter = index + flat position)."""
        if np.isscalar(shape):
            shape = (int(shape),)
        n = int(np.prod(shape)) if len(shape) else 1
        flat = BhArray((n,), np.uint64)
        if n > 0:
            api.random123(flat.base.handle, 0, (n,), (1,), self.index, self.key)
            self.index += n
        return flat.reshape(shape)

    def random_sample(self, shape):
        """Uniform float64 in [0, 1): float64(u64) / float64(2**64 - 1)."""
        return self.random123(shape).astype(np.float64) / np.iinfo(np.uint64).max

    random = random_sample

    def uniform(self, low=0.0, high=1.0, shape=None):
        return self.random_sample(shape) * (high - low) + low

    def randint(self, low, high=None, shape=None):
        if high is None:
            low, high = 0, low
        r = self.random123(shape) % np.uint64(high - low)
        return r.astype(np.int64) + low

    def standard_normal(self, shape=None, dtype=np.float64):
        """Box-Muller from two independent uniform draws."""
        u1 = 1.0 - self.random_sample(shape)
        u2 = self.random_sample(shape)
        from .bharray import ufunc
        z = ufunc("sqrt")(ufunc("log")(u1) * -2.0) * ufunc("cos")(u2 * (2.0 * np.pi))
        return z if np.dtype(dtype) == np.float64 else z.astype(dtype)

    def normal(self, loc=0.0, scale=1.0, shape=None, dtype=np.float64):
        return self.standard_normal(shape, dtype) * scale + loc

    def random_of_dtype(self, dtype=np.float64, shape=None):
        """Random values of `dtype` like the reference test helper (test/python/util.py:55-59)."""
        dtype = np.dtype(dtype)
        if dtype == np.bool_:
            return (self.random123(shape) % np.uint64(2)) == 1
        if dtype.kind in "iu":
            top = min(np.iinfo(dtype).max, 100000)
            return (self.random123(shape) % np.uint64(top)).astype(dtype)
        if dtype.kind == "c":
            re = self.random_sample(shape)
            im = self.random_sample(shape)
            return (re + im * 1j).astype(dtype)
        return self.random_sample(shape).astype(dtype)


_inst = RandomState(0)
seed = _inst.seed
random_sample = _inst.random_sample
random = _inst.random_sample
uniform = _inst.uniform
randint = _inst.randint
standard_normal = _inst.standard_normal
normal = _inst.normal
