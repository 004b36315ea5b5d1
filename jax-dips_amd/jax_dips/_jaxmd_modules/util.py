"""dtype aliases of the reference (jax_dips/_jaxmd_modules/util.py:13-21)."""
import numpy as np

Array = np.ndarray
f16 = np.float16
f32 = np.float32
f64 = np.float64
i16 = np.int16
i32 = np.int32
i64 = np.int64
