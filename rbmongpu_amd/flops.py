"""Algorithmic work of the RBM sampler, counted once from the implemented
formulas (csrc/rbm_device.hpp, csrc/rbm_kernels.hip) and used for roofline
figures.  One complex multiply is 6 FLOP, a complex add 2, a reciprocal 1.

Per hidden unit:
  RE_LOGCOSH  Re my_logcosh(z): denominator 22, numerator 3, real part of the
              quotient 8, linear term 3                         = 36
  LOGCOSH     complex my_logcosh: 22 + 3 + complex quotient 12 + 5 = 42
  TANH        my_tanh: denominator 22, numerator Horner 51, square 6,
              quotient 12                                         = 91
  PROPOSAL    theta' = theta + 2 s W (4) + Re logcosh (36) + difference and
              accumulation (2)                                    = 42
  TERM        one single-flip operator term: theta' (4) + logcosh (42) +
              difference and accumulation (4)                     = 50
Integer work (popcount parities of diagonal strings, Philox) is not counted.
"""

RE_LOGCOSH = 36
LOGCOSH = 42
TANH = 91
PROPOSAL = 4 + RE_LOGCOSH + 2
TERM_PER_FLIP = 4
TERM = LOGCOSH + 4


def sample_flops(N, M, terms, flips_per_term=1, sweeps=1, with_tanh=True):
    """FLOP of one measured sample: sweeps*N proposals, log psi, E_loc over
    the off-diagonal flip-mask terms and (for O_k) my_tanh of every angle."""
    sweep = sweeps * N * PROPOSAL * M
    measure = (LOGCOSH + 2) * M
    e_loc = terms * (TERM + TERM_PER_FLIP * flips_per_term) * M
    tanh = TANH * M if with_tanh else 0
    return sweep + measure + e_loc + tanh


def mc_launch_flops(N, M, terms, chains, samples_per_chain, therm, sweeps=1, flips_per_term=1, with_tanh=True):
    """FLOP of one Metropolis launch (thermalisation + measured samples)."""
    per_chain = therm * N * PROPOSAL * M + samples_per_chain * sample_flops(
        N, M, terms, flips_per_term, sweeps, with_tanh)
    return chains * per_chain


def mc_launch_bytes(N, M, chains, samples_per_chain, tanh_bytes=8, with_tanh=True):
    """Algorithmic HBM bytes of one launch: W and b read once per workgroup
    (16 chains), per-sample records written (spins 16 B, log psi 16 B,
    E_loc 16 B, my_tanh M x tanh_bytes)."""
    groups = (chains + 15) // 16
    reads = groups * (N * M + M) * tanh_bytes
    per_sample = 48 + (M * tanh_bytes if with_tanh else 0)
    return reads + chains * samples_per_chain * per_sample
