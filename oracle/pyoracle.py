"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports
it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

I64P = C.POINTER(C.c_int64)
DP = C.POINTER(C.c_double)
U64P = C.POINTER(C.c_uint64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run oracle/build.sh")
        L = C.CDLL(_LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_char_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_exact_energy.restype = C.c_double
        L.orc_exact_energy.argtypes = [C.c_void_p]
        L.orc_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_philox.restype = None
        for name in dir(L):
            pass
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_error().decode())


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_p(c, C.POINTER(C.c_uint32)), _p(k, C.POINTER(C.c_uint32)),
                     _p(o, C.POINTER(C.c_uint32)))
    return o


def bravais_updown(n_uc, n_uc_b, h_shift, uc, d, step):
    u = C.c_int64()
    dn = C.c_int64()
    _chk(lib().orc_bravais_updown(C.c_int64(n_uc), C.c_int64(n_uc_b), C.c_int64(h_shift),
                                  C.c_int64(uc), C.c_int64(d), C.c_int64(step),
                                  C.byref(u), C.byref(dn)))
    return u.value, dn.value


class Oracle:
    """One model (lattice + Hamiltonian + machine) built from .ini text."""

    def __init__(self, ini_text):
        self._h = lib().orc_create(ini_text.encode())
        if not self._h:
            raise RuntimeError("oracle: " + lib().orc_error().decode())
        info = np.zeros(11, dtype=np.int64)
        _chk(lib().orc_info(C.c_void_p(self._h), _p(info, I64P)))
        (self.N, self.P, self.H, self.n_symm, self.n_vb, self.n_alpha, self.n_pf,
         self.n_bonds, self.n_terms, self.n_words, self.n_neural) = [int(x) for x in info]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(C.c_void_p(self._h))
            self._h = None

    @property
    def h(self):
        return C.c_void_p(self._h)

    # ---- lattice / model tables
    def bonds(self):
        out = np.zeros(3 * self.n_bonds, dtype=np.int64)
        _chk(lib().orc_bonds(self.h, _p(out, I64P)))
        return out.reshape(-1, 3)

    def nns(self, site, n_coord):
        out = np.zeros(n_coord, dtype=np.int64)
        _chk(lib().orc_nns(self.h, C.c_int64(site), _p(out, I64P)))
        return out

    def hexagons(self):
        out = np.zeros(6 * self.N, dtype=np.int64)
        n = C.c_int64()
        _chk(lib().orc_hexagons(self.h, _p(out, I64P), C.byref(n)))
        return out[: 6 * n.value].reshape(-1, 6)

    def plaqs(self):
        out = np.zeros(5 * self.N, dtype=np.int64)
        n = C.c_int64()
        _chk(lib().orc_plaqs(self.h, _p(out, I64P), C.byref(n)))
        return out[: 5 * n.value].reshape(-1, 5)

    def symmetry(self):
        out = np.zeros(self.n_symm * self.N, dtype=np.int64)
        _chk(lib().orc_symmetry(self.h, _p(out, I64P)))
        return out.reshape(self.n_symm, self.N)

    def pf_map(self):
        """Pfaffian index map bs_ (N x N, bs[i, j]) and the number of fs_ rows"""
        out = np.zeros(self.N * self.N, dtype=np.int64)
        rows = C.c_int64()
        _chk(lib().orc_pf_map(self.h, _p(out, I64P), C.byref(rows)))
        return out.reshape(self.N, self.N).T.copy(), rows.value

    def exact_energy(self):
        return lib().orc_exact_energy(self.h)

    def terms(self):
        res = []
        for t in range(self.n_terms):
            k = C.c_int64()
            _chk(lib().orc_term_info(self.h, C.c_int64(t), C.byref(k)))
            sites = np.zeros(k.value, dtype=np.int64)
            d = 1 << k.value
            mat = np.zeros(d * d, dtype=np.complex128)
            _chk(lib().orc_term(self.h, C.c_int64(t), _p(sites, I64P), _p(mat, DP)))
            res.append((sites, mat.reshape(d, d)))
        return res

    # ---- parameters
    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.complex128)
        assert p.size == self.P
        _chk(lib().orc_set_params(self.h, _p(p, DP)))

    def get_params(self):
        p = np.zeros(self.P, dtype=np.complex128)
        _chk(lib().orc_get_params(self.h, _p(p, DP)))
        return p

    # ---- per-state quantities
    def _states(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint64).reshape(-1, self.n_words)
        return s, s.shape[0]

    def theta(self, state):
        s, _ = self._states(state)
        out = np.zeros(self.H, dtype=np.complex128)
        _chk(lib().orc_theta(self.h, _p(s, U64P), _p(out, DP)))
        return out

    def logpsi(self, states):
        s, n = self._states(states)
        out = np.zeros(n, dtype=np.complex128)
        _chk(lib().orc_logpsi(self.h, _p(s, U64P), C.c_int64(n), _p(out, DP)))
        return out

    def psi_ratio(self, state, flips):
        s, _ = self._states(state)
        f = np.ascontiguousarray(flips, dtype=np.int64)
        out = np.zeros(1, dtype=np.complex128)
        _chk(lib().orc_psi_ratio(self.h, _p(s, U64P), _p(f, I64P), C.c_int64(f.size), _p(out, DP)))
        return out[0]

    def eloc(self, states):
        s, n = self._states(states)
        out = np.zeros(n, dtype=np.complex128)
        _chk(lib().orc_eloc(self.h, _p(s, U64P), C.c_int64(n), _p(out, DP)))
        return out

    def derivative(self, states):
        """O matrix P x n (column per state)."""
        s, n = self._states(states)
        out = np.zeros(n * self.P, dtype=np.complex128)
        _chk(lib().orc_derivative(self.h, _p(s, U64P), C.c_int64(n), _p(out, DP)))
        return out.reshape(n, self.P).T

    def update_check(self, state, flips):
        s, _ = self._states(state)
        f = np.ascontiguousarray(flips, dtype=np.int64)
        d = C.c_double()
        _chk(lib().orc_update_check(self.h, _p(s, U64P), _p(f, I64P), C.c_int64(f.shape[0]),
                                    C.c_int64(f.shape[1]), C.byref(d)))
        return d.value

    def pf_context(self, state):
        s, _ = self._states(state)
        inv = np.zeros(self.N * self.N, dtype=np.complex128)
        mant = np.zeros(2)
        e = C.c_int64()
        _chk(lib().orc_pf_context(self.h, _p(s, U64P), _p(inv, DP), _p(mant, DP), C.byref(e)))
        return inv.reshape(self.N, self.N, order="F"), complex(mant[0], mant[1]), e.value

    def pf_update_check(self, state, flips):
        s, _ = self._states(state)
        f = np.ascontiguousarray(flips, dtype=np.int64)
        a = C.c_double()
        b = C.c_double()
        ratios = np.zeros(f.shape[0], dtype=np.complex128)
        _chk(lib().orc_pf_update_check(self.h, _p(s, U64P), _p(f, I64P), C.c_int64(f.shape[0]),
                                       C.c_int64(f.shape[1]), C.byref(a), C.byref(b), _p(ratios, DP)))
        return a.value, b.value, ratios

    # ---- sampling / statistics / SR
    def sample(self, seed, epoch, n_samples, n_chains, threads=0, use_lut=False, want_O=False):
        st = np.zeros((n_samples, self.n_words), dtype=np.uint64)
        e = np.zeros(n_samples, dtype=np.complex128)
        O = np.zeros(n_samples * self.P, dtype=np.complex128) if want_O else None
        acc = np.zeros(n_chains)
        _chk(lib().orc_sample(self.h, C.c_uint64(seed), C.c_uint32(epoch), C.c_int(threads),
                              C.c_int(int(use_lut)), _p(st, U64P), _p(e, DP),
                              _p(O, DP) if O is not None else None, _p(acc, DP)))
        if O is not None:
            O = O.reshape(n_samples, self.P).T
        return st, e, O, acc

    def aggregate(self, e, O=None, n_bins=0):
        e = np.ascontiguousarray(e, dtype=np.complex128)
        S = e.size
        oe = np.zeros(6)
        oo = np.zeros(6 * self.P)
        Ot = None
        if O is not None:
            Ot = np.ascontiguousarray(np.asarray(O).T, dtype=np.complex128)
        _chk(lib().orc_aggregate(self.h, _p(e, DP), _p(Ot, DP) if Ot is not None else None,
                                 C.c_int64(S), C.c_int64(n_bins), _p(oe, DP), _p(oo, DP)))
        P = self.P
        res = dict(mean=complex(oe[0], oe[1]), var=oe[2], var_binned=oe[3], tau=oe[4], std=oe[5])
        if O is not None:
            res["o_mean"] = oo[: 2 * P].view(np.complex128).copy()
            res["o_var"] = oo[2 * P: 3 * P].copy()
            res["eo"] = oo[3 * P: 5 * P].view(np.complex128).copy()
        return res

    def epoch(self, seed, epoch, threads=0, use_lut=False):
        st = np.zeros(10)
        _chk(lib().orc_epoch(self.h, C.c_uint64(seed), C.c_uint32(epoch), C.c_int(threads),
                             C.c_int(int(use_lut)), _p(st, DP)))
        keys = ["energy", "std", "acc", "tau", "itn", "rs", "max_dw", "ok", "t_sample", "t_opt"]
        return dict(zip(keys, st.tolist()))

    def exact_expectation(self):
        out = np.zeros(2)
        _chk(lib().orc_exact_expectation(self.h, _p(out, DP)))
        return complex(out[0], out[1])

    def clusters(self):
        """cluster catalogue of the learned proposal: cyc (N x kmax)"""
        cyc = np.zeros(self.N * 8, dtype=np.int64)
        k = C.c_int64()
        _chk(lib().orc_clusters(self.h, _p(cyc, I64P), C.byref(k)))
        return cyc[: self.N * k.value].reshape(self.N, k.value)


def _mlp_args(W1, b1, W2, b2):
    W1 = np.ascontiguousarray(W1, dtype=np.float64)
    W2 = np.ascontiguousarray(W2, dtype=np.float64)
    return (W1, np.ascontiguousarray(b1, dtype=np.float64), W2, np.ascontiguousarray(b2, dtype=np.float64))


def mlp_logits(W1, b1, W2, b2, states, N):
    """FP64 logits z (n x N) of the proposal network for packed states"""
    W1, b1, W2, b2 = _mlp_args(W1, b1, W2, b2)
    Hn = W1.shape[0]
    st = np.ascontiguousarray(states, dtype=np.uint64).reshape(len(states), -1)
    z = np.zeros((len(st), N))
    _chk(lib().orc_mlp_logits(C.c_int64(N), C.c_int64(Hn), _p(W1, DP), _p(b1, DP), _p(W2, DP), _p(b2, DP),
                              _p(st, U64P), C.c_int64(st.shape[1]), C.c_int64(len(st)), _p(z, DP)))
    return z


def mlp_grad(W1, b1, W2, b2, states, N, seeds, w, beta):
    """loss and gradients (gW1, gb1, gW2, gb2) of the proposal-network
    batch objective"""
    W1, b1, W2, b2 = _mlp_args(W1, b1, W2, b2)
    Hn = W1.shape[0]
    st = np.ascontiguousarray(states, dtype=np.uint64).reshape(len(states), -1)
    sd = np.ascontiguousarray(seeds, dtype=np.int64)
    ww = np.ascontiguousarray(w, dtype=np.float64)
    g = [np.zeros_like(W1), np.zeros_like(b1), np.zeros_like(W2), np.zeros_like(b2)]
    loss = C.c_double()
    _chk(lib().orc_mlp_grad(C.c_int64(N), C.c_int64(Hn), _p(W1, DP), _p(b1, DP), _p(W2, DP), _p(b2, DP),
                            _p(st, U64P), C.c_int64(st.shape[1]), C.c_int64(len(st)), _p(sd, I64P), _p(ww, DP),
                            C.c_double(beta), *[_p(x, DP) for x in g], C.byref(loss)))
    return loss.value, g


def apply_S(O, mean, var, r1, r2, rd, n_neural, x):
    P, S = O.shape
    Ot = np.ascontiguousarray(O.T, dtype=np.complex128)
    mean = np.ascontiguousarray(mean, dtype=np.complex128)
    var = np.ascontiguousarray(var, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.zeros(P, dtype=np.complex128)
    _chk(lib().orc_apply_S(C.c_int64(P), C.c_int64(S), C.c_int64(n_neural), _p(Ot, DP), _p(mean, DP),
                           _p(var, DP), C.c_double(r1), C.c_double(r2), C.c_double(rd), _p(x, DP),
                           _p(y, DP)))
    return y


def cg(O, mean, var, r1, r2, rd, n_neural, b, x0, max_iter, rtol):
    P, S = O.shape
    Ot = np.ascontiguousarray(O.T, dtype=np.complex128)
    mean = np.ascontiguousarray(mean, dtype=np.complex128)
    var = np.ascontiguousarray(var, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.array(x0, dtype=np.complex128)
    itn = C.c_int64()
    rs = C.c_double()
    _chk(lib().orc_cg(C.c_int64(P), C.c_int64(S), C.c_int64(n_neural), _p(Ot, DP), _p(mean, DP),
                      _p(var, DP), C.c_double(r1), C.c_double(r2), C.c_double(rd), _p(b, DP), _p(x, DP),
                      C.c_int64(max_iter), C.c_double(rtol), C.byref(itn), C.byref(rs)))
    return x, itn.value, rs.value


def random_states(rng, n, N):
    """n random states packed as uint64 words (bit i <=> sigma_i = -1)."""
    nw = (N + 63) // 64
    bits = rng.integers(0, 2, size=(n, N), dtype=np.uint64)
    out = np.zeros((n, nw), dtype=np.uint64)
    for i in range(N):
        out[:, i // 64] |= bits[:, i] << np.uint64(i % 64)
    return out


def sigma_of(states, N):
    """+-1 spin values of packed states, shape (n, N)."""
    s = np.asarray(states, dtype=np.uint64).reshape(-1, (N + 63) // 64)
    out = np.empty((s.shape[0], N))
    for i in range(N):
        out[:, i] = 1.0 - 2.0 * ((s[:, i // 64] >> np.uint64(i % 64)) & np.uint64(1)).astype(np.float64)
    return out


# ---- optimizer plugins and SGD gradient (numpy restatement; O(P) vector
# arithmetic, so numpy is the oracle here)
def adam_steps(dws, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """adam_plugin::apply (plugin.cpp:43-57) over a sequence of solver steps"""
    n = len(dws[0])
    m = np.zeros(n, np.complex128)
    wr = np.zeros(n)
    wi = np.zeros(n)
    out = []
    for t, dw in enumerate(dws, start=1):
        dw = np.asarray(dw, np.complex128)
        m = beta1 * m + (1 - beta1) * dw
        wr = beta2 * wr + (1 - beta2) * dw.real ** 2
        wi = beta2 * wi + (1 - beta2) * dw.imag ** 2
        mh = m / (1 - beta1 ** t)
        wrh = wr / (1 - beta2 ** t)
        wih = wi / (1 - beta2 ** t)
        out.append(lr * (mh.real / (np.sqrt(wrh) + eps) + 1j * mh.imag / (np.sqrt(wih) + eps)))
    return out


def momentum_steps(dws, lr, alpha=0.9, dialup=1.0):
    """momentum_plugin::apply (plugin.cpp:69-77) over a sequence of steps"""
    m = np.zeros(len(dws[0]), np.complex128)
    dup = 1.0 if dialup == 1.0 else 1e-3
    out = []
    for dw in dws:
        dw = np.asarray(dw, np.complex128)
        ax = alpha
        if np.all(np.abs(m) <= 1e-12):  # Eigen isZero
            m = dw.copy()
        if dup < 1.0:
            dup *= dialup
            ax *= dup
        m = ax * m + (1 - ax) * dw
        out.append(lr * m)
    return out
