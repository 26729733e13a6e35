"""JSON encoding of numpy arrays in the reference's on-disk format
(pyRBMonGPU/json_numpy.py:5-38): {"type": "ndarray", "dtype", "complex",
"real"/"imag" or "data"}."""
import json

import numpy as np


class NumpyEncoder(json.JSONEncoder):
    def default(self, obj):
        if isinstance(obj, np.ndarray):
            out = {"type": "ndarray", "dtype": str(obj.dtype), "complex": bool(np.iscomplexobj(obj))}
            if out["complex"]:
                out["real"] = obj.real.tolist()
                out["imag"] = obj.imag.tolist()
            else:
                out["data"] = obj.tolist()
            return out
        if isinstance(obj, (np.integer,)):
            return int(obj)
        if isinstance(obj, (np.floating,)):
            return float(obj)
        return super().default(obj)


class NumpyDecoder(json.JSONDecoder):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, object_hook=self._hook, **kwargs)

    @staticmethod
    def _hook(obj):
        if obj.get("type") == "ndarray":
            dtype = np.dtype(obj["dtype"])
            if obj["complex"]:
                return np.array(obj["real"], dtype=dtype) + 1j * np.array(obj["imag"], dtype=dtype)
            return np.array(obj["data"], dtype=dtype)
        return obj
