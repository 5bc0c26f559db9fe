"""The radix-8 stage measurement (mpa_bench_dft8): both variants compute the
8-point forward DFT.

The vector-ALU butterflies (the FFT passes' own dft<8>) and the matrix-core
form (a real 16 x 16 product on v_mfma_f32_16x16x4f32, which also pins that
instruction's operand lane layout) against numpy's FFT, applied 1 and 3 times.
"""

import numpy as np
import pytest

from conftest import rel_max_err

from mumaxplus import _mumaxpluscpp as cpp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("rounds", [1, 3])
def test_dft8_stage_variants_match_numpy(variant, rounds):
    rng = np.random.default_rng(7 + variant + 10 * rounds)
    lines = (rng.normal(size=(1024, 8)) + 1j * rng.normal(size=(1024, 8))).astype(np.complex64)
    got, ms = cpp._bench_dft8(variant, lines, rounds)
    want = lines.astype(np.complex128)
    for _ in range(rounds):
        want = np.fft.fft(want, axis=1)
    assert ms > 0
    assert rel_max_err(got.view(np.float32), want.astype(np.complex64).view(np.float32)) < 2e-6 * rounds


def test_dft8_bench_refuses_bad_sizes():
    lines = np.zeros((100, 8), dtype=np.complex64)
    with pytest.raises(ValueError, match="multiple of 256"):
        cpp._bench_dft8(0, lines, 1)
    with pytest.raises(ValueError, match="variant"):
        cpp._bench_dft8(2, np.zeros((256, 8), dtype=np.complex64), 1)
