"""ctypes binding of the warm-start MLP (include/cumf_mlp.h)."""
import ctypes as C

import numpy as np

from .als import CumfError, _check, _ptr, lib

P = C.c_void_p


class MLP:
    """(2f+1) -> 256 -> 256 -> f network on one GPU; FP32 master weights."""

    def __init__(self, f, seed=0, device=0):
        self.f = int(f)
        self._h = P()
        _check(lib().cumf_mlp_create(device, self.f, int(seed), C.byref(self._h)), "cumf_mlp_create")
        dims = np.zeros(4, np.int32)
        cnt = C.c_int64()
        _check(lib().cumf_mlp_dims(self._h, _ptr(dims), C.byref(cnt)), "cumf_mlp_dims")
        self.k0, self.h1, self.h2, self.n3 = (int(v) for v in dims)
        self.nparams = int(cnt.value)

    def close(self):
        if self._h:
            lib().cumf_mlp_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # Flat parameter vector <-> [W1, c1, W2, c2, W3, c3] (W: out x in).
    def shapes(self):
        return [(self.h1, self.k0), (self.h1,), (self.h2, self.h1), (self.h2,), (self.n3, self.h2), (self.n3,)]

    def split(self, flat):
        out, o = [], 0
        for sh in self.shapes():
            n = int(np.prod(sh))
            out.append(flat[o:o + n].reshape(sh))
            o += n
        return out

    def get_params(self):
        flat = np.empty(self.nparams, np.float32)
        _check(lib().cumf_mlp_get_params(self._h, _ptr(flat)), "cumf_mlp_get_params")
        return self.split(flat)

    def set_params(self, params):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).ravel() for p in params]))
        if flat.shape[0] != self.nparams:
            raise CumfError("parameter count mismatch")
        _check(lib().cumf_mlp_set_params(self._h, _ptr(flat)), "cumf_mlp_set_params")

    def _feats(self, X):
        X = np.ascontiguousarray(X, np.float32)
        assert X.ndim == 2 and X.shape[1] == self.k0
        return X

    def predict(self, X):
        X = self._feats(X)
        out = np.empty((X.shape[0], self.f), np.float32)
        _check(lib().cumf_mlp_predict_host(self._h, _ptr(X), X.shape[0], _ptr(out)), "cumf_mlp_predict_host")
        return out

    def gradients(self, X, T):
        X = self._feats(X)
        T = np.ascontiguousarray(T[:, :self.f], np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _check(lib().cumf_mlp_gradients_host(self._h, _ptr(X), _ptr(T), X.shape[0], _ptr(g), C.byref(loss)),
               "cumf_mlp_gradients_host")
        return self.split(g), loss.value

    def step(self, X, T, lr):
        X = self._feats(X)
        T = np.ascontiguousarray(T[:, :self.f], np.float32)
        loss = C.c_double()
        _check(lib().cumf_mlp_step_host(self._h, _ptr(X), _ptr(T), X.shape[0], float(lr), C.byref(loss)),
               "cumf_mlp_step_host")
        return loss.value

    def harvest(self, als, iters):
        _check(lib().cumf_mlp_harvest(als._h, self._h, int(iters)), "cumf_mlp_harvest")

    def harvested(self):
        return int(lib().cumf_mlp_harvested(self._h))

    def clear_harvest(self):
        _check(lib().cumf_mlp_clear_harvest(self._h), "cumf_mlp_clear_harvest")

    def train(self, epochs, batch=4096, lr=1e-3, seed=0):
        loss = C.c_double()
        _check(lib().cumf_mlp_train(self._h, int(epochs), int(batch), float(lr), int(seed), C.byref(loss)),
               "cumf_mlp_train")
        return loss.value
