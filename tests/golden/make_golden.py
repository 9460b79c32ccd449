"""Writes tests/golden/pcg32_streams.npz.

The words come from an independent pure-Python PCG32 (O'Neill 2014, XSH-RR,
multiplier 6364136223846793005) seeded as Pcg32::new(idx, seed) does
(renderer/rng.rs:16-26: inc = ((idx+1) << 1) | 1; step; state +=
0x853c49e6748fea9b + seed; step).  The oracle and the device code must
reproduce them bit for bit.
"""
import os

import numpy as np

M64 = (1 << 64) - 1


def pcg32_words(idx, seed, n):
    state, inc = 0, (((idx + 1) << 1) | 1) & M64

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + (inc | 1)) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + ((0x853c49e6748fea9b + seed) & M64)) & M64
    step()
    return [step() for _ in range(n)]


def main():
    cases = [(0, 0), (1, 0), (1 << 20, 0), ((1 << 32) - 1, 0), (12345, 7), (4194303, 0)]
    words = np.array([pcg32_words(i, s, 8) for i, s in cases], np.uint32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg32_streams.npz")
    np.savez(out, idx=np.array([c[0] for c in cases], np.uint64), seed=np.array([c[1] for c in cases], np.uint64),
             words=words)
    print("wrote", out)


if __name__ == "__main__":
    main()
