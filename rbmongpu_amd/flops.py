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


def mc_launch_flops(N, M, chains, samples_per_chain, therm, sweeps=1):
    """FLOP of one Metropolis launch: (therm + samples * sweeps) * N proposals
    per chain, PROPOSAL FLOP per hidden unit."""
    return chains * (therm + samples_per_chain * sweeps) * N * PROPOSAL * M


def eval_flops(N, M, terms, n, flips_per_term=1, with_tanh=True):
    """FLOP of evaluating n configurations: angles from scratch (4 N M),
    log psi, E_loc over the off-diagonal flip-mask terms and my_tanh."""
    per = 4 * N * M + (LOGCOSH + 2) * M + terms * (TERM + TERM_PER_FLIP * flips_per_term) * M
    if with_tanh:
        per += TANH * M
    return n * per


def mc_launch_bytes(chains, samples_per_chain):
    """Algorithmic HBM bytes of one Metropolis launch: one 16-byte spin
    record per sample (W is read into LDS once per workgroup)."""
    return chains * samples_per_chain * 16


def eval_bytes(M, n, tanh_bytes=8, with_tanh=True):
    """Algorithmic HBM bytes of evaluating n recorded samples: read 16-byte
    configurations, write log psi and E_loc (16 B each) and M my_tanh values."""
    return n * (16 + 32 + (M * tanh_bytes if with_tanh else 0))


def deep_forward_mfma_flops(N, sizes):
    """Matrix-core FLOP of one dense PsiDeep forward pass: 4 in out for the
    first layer (real spins times complex weights) and 8 in out for every
    later complex layer (four real products)."""
    flop, prev = 0, N
    for l, size in enumerate(sizes):
        flop += (4 if l == 0 else 8) * prev * size
        prev = size
    return flop


def deep_forward_valu_flops(sizes):
    """Vector FLOP of one forward pass: complex my_logcosh per unit and the
    final weighted sum (6 + 2 per unit of the last layer)."""
    return LOGCOSH * sum(sizes) + 8 * sizes[-1]


def deep_mc_launch_flops(N, sizes, chains, proposals_per_chain):
    """Matrix-core FLOP of one PsiDeep Metropolis launch: one full forward pass
    per proposal and one at the start of each chain."""
    return chains * (proposals_per_chain + 1) * deep_forward_mfma_flops(N, sizes)


def s_matrix_flops(P, samples):
    """FP32 matrix-core FLOP of the centred S-matrix Gram product on real rows
    of 2P values per sample: the symmetric product needs its lower triangle,
    2 (2P)(2P + 1) / 2 FLOP per sample."""
    return float(2 * P) * (2 * P + 1) * samples


def cholesky_flops(n):
    """Real FLOP of a complex Hermitian Cholesky factorisation A = L L^H:
    the updates of the lower triangle are n^3 / 6 complex multiply-adds of
    8 real FLOP (LAPACK's zpotrf count, 4 n^3 / 3 to leading order)."""
    return 4.0 * n ** 3 / 3.0
