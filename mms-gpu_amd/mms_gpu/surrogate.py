"""Training of the learned far-field surrogate (wraps ``mms_sur_trainer``).

The trained weights load into a film with ``Film.surrogate_init(weights)``.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib


def surrogate_state(xdim, ydim, zdim, Ms, seed, index):
    """Film state `index` of the seeded training family: (M float32 [n*3],
    material uint8 [n], kind = 2 * shape + state)."""
    n = xdim * ydim * zdim
    M = np.zeros(n * 3, np.float32)
    mat = np.zeros(n, np.uint8)
    kind = C.c_int(-1)
    check(lib().mms_surrogate_state(xdim, ydim, zdim, float(Ms), int(seed), int(index), M,
                                    mat.ctypes.data_as(C.c_void_p), C.byref(kind)), "mms_surrogate_state")
    return M, mat, kind.value


class SurrogateTrainer:
    """Adam training of the surrogate on (charge, potential) pairs of films of
    xdim x ydim x zdim cubic cells of side dx, `batch` films per step."""

    def __init__(self, xdim, ydim, zdim, dx, Ms, batch, weights=None):
        self.dims = (xdim, ydim, zdim)
        self.n = xdim * ydim * zdim
        self.batch = batch
        self._t = C.c_void_p()
        w = None
        if weights is not None:
            self._w = np.ascontiguousarray(weights, np.float32)
            if self._w.size != lib().mms_surrogate_weight_count():
                raise ValueError("surrogate weight vector has the wrong size")
            w = self._w.ctypes.data_as(C.c_void_p)
        check(lib().mms_surrogate_trainer_create(C.byref(self._t), xdim, ydim, zdim, float(dx), float(Ms),
                                                 batch, w), "mms_surrogate_trainer_create")

    def close(self):
        if self._t:
            lib().mms_surrogate_trainer_destroy(self._t)
            self._t = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _idx(self, idx):
        idx = np.ascontiguousarray(idx, np.int64)
        if idx.size != self.batch:
            raise ValueError(f"a batch holds {self.batch} pair indices")
        return idx

    def generate(self, npairs, seed, P=6):
        """Append npairs pairs of the seeded film states (potential by FMM of
        order P; P <= 0: the direct sum)."""
        check(lib().mms_surrogate_trainer_generate(self._t, int(npairs), int(seed), int(P)),
              "mms_surrogate_trainer_generate")

    def add(self, charge, potential):
        charge = np.ascontiguousarray(charge, np.float32).reshape(-1, self.n)
        potential = np.ascontiguousarray(potential, np.float32).reshape(-1, self.n)
        check(lib().mms_surrogate_trainer_add(self._t, charge.shape[0], charge.ravel(), potential.ravel()),
              "mms_surrogate_trainer_add")

    @property
    def pairs(self):
        return int(lib().mms_surrogate_trainer_pairs(self._t))

    def pair(self, i):
        rho = np.zeros(self.n, np.float32)
        phi = np.zeros(self.n, np.float32)
        check(lib().mms_surrogate_trainer_get_pair(self._t, int(i), rho.ctypes.data_as(C.c_void_p),
                                                   phi.ctypes.data_as(C.c_void_p)), "mms_surrogate_trainer_get_pair")
        return rho, phi

    def gradient(self, idx):
        """(gradient in API weight order, loss) of the batch idx, no update."""
        g = np.zeros(lib().mms_surrogate_weight_count(), np.float32)
        loss = C.c_double()
        check(lib().mms_surrogate_trainer_gradient(self._t, self._idx(idx), g.ctypes.data_as(C.c_void_p),
                                                   C.byref(loss)), "mms_surrogate_trainer_gradient")
        return g, loss.value

    def step(self, idx, lr):
        loss = C.c_double()
        check(lib().mms_surrogate_trainer_step(self._t, self._idx(idx), float(lr), C.byref(loss)),
              "mms_surrogate_trainer_step")
        return loss.value

    def eval(self, idx):
        r = np.zeros(self.batch, np.float64)
        check(lib().mms_surrogate_trainer_eval(self._t, self._idx(idx), r), "mms_surrogate_trainer_eval")
        return r

    def predict(self, idx):
        out = np.zeros(self.batch * self.n, np.float32)
        check(lib().mms_surrogate_trainer_predict(self._t, self._idx(idx), out), "mms_surrogate_trainer_predict")
        return out.reshape(self.batch, self.n)

    def weights(self):
        w = np.zeros(lib().mms_surrogate_weight_count(), np.float32)
        check(lib().mms_surrogate_trainer_weights(self._t, w), "mms_surrogate_trainer_weights")
        return w
