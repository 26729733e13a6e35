"""rbmongpu_amd: RBMonGPU's variational-Monte-Carlo hot path on AMD Instinct
MI355X (gfx950).

Drop-in for the reference package pyRBMonGPU (pyRBMonGPU/__init__.py:1-26):
the same classes (Psi, PsiDeep, Operator, Spins, MonteCarloLoop,
ExactSummation, ExpectationValue, HilbertSpaceDistance) and functions
(psi_O_k_vector, get_S_matrix, activation_function, setDevice, ...), backed by
hand-written HIP kernels in librbmhip.so.  The compiled module is required:
there is no CPU fallback.

Beyond the reference: MADE, a masked autoregressive proposal network trained
on the GPU, and LearnedMonteCarloLoop, an independent Metropolis-Hastings
ensemble that proposes from it and is accepted with the exact |psi|^2 ratio.
"""
from __future__ import annotations

import json as _json

import numpy as _np

try:
    from ._pyRBMonGPU import (  # noqa: F401
        Psi, PsiDeep, Operator, Spins, MonteCarloLoop, ExactSummation, ExpectationValue,
        HilbertSpaceDistance, psi_O_k_vector, get_S_matrix, activation_function, activation_derivative,
        setDevice, start_profiling, stop_profiling, device_count, timing_enable, timing_reset, timing_query,
        log_psi_batch, local_energy_batch, synchronize, MADE, LearnedMonteCarloLoop,
        get_S_matrix_centred, sr_update, cholesky_solve,
    )
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "rbmongpu_amd: the compiled HIP module _pyRBMonGPU is missing; build it with "
        "`python rbmongpu_amd/build.py` (needs hipcc for gfx950)") from exc

from .json_numpy import NumpyEncoder, NumpyDecoder  # noqa: E402
from .new_neural_network import new_neural_network, new_deep_neural_network, complex_noise, real_noise  # noqa: E402,F401
from . import hamiltonians  # noqa: E402,F401
from . import gradient_descent  # noqa: E402,F401
from .gradient_descent import *  # noqa: E402,F401,F403
from .learning import LearningByGradientDescent, DidNotConverge, L2Regularization  # noqa: E402,F401


def _normalize_psi(self, exact_summation):
    """Psi.normalize: prefactor /= norm (pyRBMonGPU/Psi.py:54-55)."""
    self.prefactor /= self.norm(exact_summation)


def _normalize_deep(self, exact_summation):
    """PsiDeep.normalize: prefactor = 1 / norm (pyRBMonGPU/PsiDeep.py:55-57)."""
    self.prefactor = 1
    self.prefactor /= self.norm(exact_summation)


def _psi_to_json(self):
    """Psi.to_json (pyRBMonGPU/Psi.py:8-21)."""
    obj = dict(type="Psi", alpha=self.alpha, beta=self.beta, b=self.b, W=self.W, prefactor=self.prefactor)
    return _json.loads(_json.dumps(obj, cls=NumpyEncoder))


def _load_array(obj):
    if isinstance(obj, _np.ndarray):
        return obj
    return _json.loads(_json.dumps(obj), cls=NumpyDecoder)


@staticmethod
def _psi_from_json(json_obj, gpu):
    """Psi.from_json (pyRBMonGPU/Psi.py:36-45), with free_quantum_axis=False
    (the reference passes one argument too few to the constructor)."""
    return Psi(_load_array(json_obj["alpha"]), _load_array(json_obj["beta"]), _load_array(json_obj["b"]),
               _load_array(json_obj["W"]), json_obj["prefactor"], False, gpu)


def _deep_to_json(self):
    """PsiDeep.to_json (pyRBMonGPU/PsiDeep.py:8-23)."""
    obj = dict(type="PsiDeep", alpha=self.alpha, beta=self.beta, b=self.b, connections=self.connections, W=self.W,
               final_weights=self.final_weights, prefactor=self.prefactor, free_quantum_axis=self.free_quantum_axis)
    return _json.loads(_json.dumps(obj, cls=NumpyEncoder))


@staticmethod
def _deep_from_json(json_obj, gpu):
    """PsiDeep.from_json (pyRBMonGPU/PsiDeep.py:26-46)."""
    obj = _json.loads(_json.dumps(json_obj, cls=NumpyEncoder), cls=NumpyDecoder)
    final = obj["final_weights"] if "final_weights" in obj else _np.ones(_np.array(obj["W"][-1]).shape[1])
    return PsiDeep(obj["alpha"], obj["beta"], list(obj["b"]), [_np.asarray(c, dtype=_np.uint32) for c in obj["connections"]],
                   list(obj["W"]), final, obj["prefactor"], obj["free_quantum_axis"], gpu)


def _first_layer_params_slice(self):
    """PsiDeep.first_layer_params_slice (pyRBMonGPU/PsiDeep.py:79-81)."""
    return slice(2 * self.N, 2 * self.N + self.W[0].size + self.b[0].size)


def _rbm_params_slice(self):
    """Psi has one layer: b, W (and a) -- the parameters the driver's L2
    regularisation acts on (the reference defines the slice for PsiDeep only)."""
    return slice(2 * self.N, self.num_params)


setattr(Psi, "normalize", _normalize_psi)
setattr(Psi, "to_json", _psi_to_json)
setattr(Psi, "from_json", _psi_from_json)
setattr(PsiDeep, "normalize", _normalize_deep)
setattr(PsiDeep, "to_json", _deep_to_json)
setattr(PsiDeep, "from_json", _deep_from_json)
setattr(PsiDeep, "first_layer_params_slice", property(_first_layer_params_slice))
setattr(Psi, "first_layer_params_slice", property(_rbm_params_slice))
