"""ctypes binding of the MI355X VMC library (include/rbmgpu.h, include/rbmhost.h).

`Model` builds lattice, Hamiltonian and machine tables from the reference's
.ini text; `Gpu` owns one device context.  The native libraries are loaded
from rbm-nqs_amd/lib; there is no fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(os.path.dirname(_PKG)), "lib")
_gpu = None
_host = None

DP = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)

FP32, FP64 = 32, 64


class SrParams(C.Structure):
    _fields_ = [("lr", C.c_double), ("reg1", C.c_double), ("reg2", C.c_double), ("reg1_delta", C.c_double),
                ("max_iter", C.c_int64), ("rtol", C.c_double), ("graph", C.c_int32)]


class EpochStats(C.Structure):
    _fields_ = [("e_mean", C.c_double * 2), ("e_var", C.c_double), ("e_var_binned", C.c_double),
                ("e_tau", C.c_double), ("e_std", C.c_double), ("n_bins", C.c_int64), ("acceptance", C.c_double),
                ("cg_itn", C.c_int64), ("cg_rs", C.c_double), ("max_dw", C.c_double), ("updated", C.c_int32),
                ("graph_replayed", C.c_int32), ("ms", C.c_double), ("explicit_s", C.c_int32), ("pad_", C.c_int32)]


class Moments(C.Structure):
    _fields_ = [("e_mean", C.c_double * 2), ("e_var", C.c_double), ("e_var_binned", C.c_double),
                ("e_tau", C.c_double), ("e_std", C.c_double), ("n_bins", C.c_int64),
                ("acceptance", C.c_double)]


def libs():
    """Load librbmgpu.so and librbmhost.so (raises if they are not built)."""
    global _gpu, _host
    if _gpu is None:
        g = os.path.join(LIBDIR, "librbmgpu.so")
        h = os.path.join(LIBDIR, "librbmhost.so")
        if not (os.path.exists(g) and os.path.exists(h)):
            raise RuntimeError(f"native libraries missing in {LIBDIR}: run rbm-nqs_amd/build.py")
        _gpu = C.CDLL(g, mode=C.RTLD_GLOBAL)
        _host = C.CDLL(h)
        _gpu.rbmgpu_last_error.restype = C.c_char_p
        _gpu.rbmgpu_last_error.argtypes = [C.c_void_p]
        _gpu.rbmgpu_destroy.argtypes = [C.c_void_p]
        _gpu.rbmgpu_destroy.restype = None
        _host.rbmhost_create.restype = C.c_void_p
        _host.rbmhost_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int32]
        _host.rbmhost_destroy.argtypes = [C.c_void_p]
        _host.rbmhost_destroy.restype = None
        _host.rbmhost_exact_energy.restype = C.c_double
        _host.rbmhost_exact_energy.argtypes = [C.c_void_p]
    return _gpu, _host


def exported_symbols():
    """Names the two libraries must export (checked by the CPU tests)."""
    return {
        "gpu": ["rbmgpu_create", "rbmgpu_destroy", "rbmgpu_last_error", "rbmgpu_n_params",
                "rbmgpu_launch_shape", "rbmgpu_n_chains", "rbmgpu_set_params", "rbmgpu_get_params", "rbmgpu_time_sx",
                "rbmgpu_learn_train", "rbmgpu_learn_grad", "rbmgpu_learn_weights", "rbmgpu_learn_logits",
                "rbmgpu_learn_set_fraction",
                "rbmgpu_update_weights", "rbmgpu_sample_epoch", "rbmgpu_load_samples",
                "rbmgpu_reduce", "rbmgpu_sr_solve", "rbmgpu_apply_S", "rbmgpu_eval_logpsi",
                "rbmgpu_psi_ratio", "rbmgpu_get_samples", "rbmgpu_n_samples",
                "rbmgpu_kernel_times", "rbmgpu_synchronize", "rbmgpu_apply_S_noalloc", "rbmgpu_train_epoch",
                "rbmgpu_sr_solve_ex", "rbmgpu_build_S", "rbmgpu_full_epoch", "rbmgpu_full_weights"],
        "host": ["rbmhost_create", "rbmhost_destroy", "rbmhost_info", "rbmhost_bonds",
                 "rbmhost_perms", "rbmhost_terms", "rbmhost_exact_energy", "rbmhost_use_evaluation",
                 "rbmhost_set_n_samples", "rbmhost_initial_weights", "rbmhost_gpu_create",
                 "rbmhost_clusters", "rbmhost_step_ok", "rbmhost_pf_map", "rbmhost_plugin_create",
                 "rbmhost_plugin_apply", "rbmhost_plugin_destroy"],
    }


def step_ok(dw):
    """massive-step guard of the driver (abstract_optimizer.cpp:157): (ok, max |dw|)"""
    dw = np.ascontiguousarray(dw, dtype=np.complex128)
    mx = C.c_double()
    rc = libs()[1].rbmhost_step_ok(dw.ctypes.data_as(DP), C.c_int64(dw.size), C.byref(mx))
    if rc < 0:
        raise RuntimeError("rbmhost_step_ok: bad arguments")
    return rc == 1, mx.value


class Plugin:
    """optimizer.plugin of the host layer (plugin.cpp:30-77): "adam" (beta1,
    beta2, eps) or "momentum" (alpha, dialup)"""

    def __init__(self, kind, n, p1=0.0, p2=0.0, p3=0.0):
        h = libs()[1]
        h.rbmhost_plugin_create.restype = C.c_void_p
        h.rbmhost_plugin_create.argtypes = [C.c_char_p, C.c_int64, C.c_double, C.c_double, C.c_double]
        h.rbmhost_plugin_destroy.argtypes = [C.c_void_p]
        h.rbmhost_plugin_destroy.restype = None
        self._h = h.rbmhost_plugin_create(kind.encode(), n, p1, p2, p3)
        if not self._h:
            raise RuntimeError(f"rbmhost: unknown plugin '{kind}'")
        self.n = n

    def apply(self, dw, lr):
        d = np.array(dw, dtype=np.complex128)
        assert d.size == self.n
        rc = libs()[1].rbmhost_plugin_apply(C.c_void_p(self._h), d.ctypes.data_as(DP), C.c_int64(d.size),
                                            C.c_double(lr))
        if rc != 0:
            raise RuntimeError("rbmhost_plugin_apply failed")
        return d

    def __del__(self):
        if getattr(self, "_h", None):
            libs()[1].rbmhost_plugin_destroy(C.c_void_p(self._h))
            self._h = None


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Model:
    """Host tables of one .ini configuration (rbmhost_*)."""

    def __init__(self, ini_text):
        _, h = libs()
        err = C.create_string_buffer(512)
        self._h = h.rbmhost_create(ini_text.encode(), err, 512)
        if not self._h:
            raise RuntimeError("rbmhost: " + err.value.decode())
        self._refresh()

    def _refresh(self):
        _, h = libs()
        o = np.zeros(17, dtype=np.int64)
        h.rbmhost_info(C.c_void_p(self._h), _p(o, I64P))
        (self.N, self.P, self.n_neural, self.n_vb, self.n_alpha, self.n_symm, self.pf_rows,
         self.n_terms, self.n_groups, self.n_bonds, self.n_chains, self.n_samples, self.n_warmup,
         self.step_size, self.n_prop_bonds, self.learn_hidden, self.cluster_kmax) = [int(x) for x in o]
        self.n_words = (self.N + 63) // 64
        self.H = self.n_alpha * self.n_symm

    def __del__(self):
        if getattr(self, "_h", None):
            libs()[1].rbmhost_destroy(C.c_void_p(self._h))
            self._h = None

    @property
    def h(self):
        return C.c_void_p(self._h)

    def bonds(self):
        out = np.zeros(3 * self.n_bonds, dtype=np.int64)
        libs()[1].rbmhost_bonds(self.h, _p(out, I64P))
        return out.reshape(-1, 3)

    def perms(self):
        out = np.zeros(self.n_symm * self.N, dtype=np.int64)
        libs()[1].rbmhost_perms(self.h, _p(out, I64P))
        return out.reshape(self.n_symm, self.N)

    def pf_map(self):
        """Pfaffian index map bs_ (N x N, bs[i, j]) and the number of fs_ rows"""
        out = np.zeros(self.N * self.N, dtype=np.int64)
        rows = C.c_int64()
        if libs()[1].rbmhost_pf_map(self.h, _p(out, I64P), C.byref(rows)) != 0:
            raise RuntimeError("rbmhost_pf_map: no Pfaffian factor")
        return out.reshape(self.N, self.N).T.copy(), rows.value

    def clusters(self):
        """cluster catalogue of the learned proposal (N x kmax)"""
        out = np.zeros(8 * self.N, dtype=np.int64)
        k = C.c_int64()
        libs()[1].rbmhost_clusters(self.h, _p(out, I64P), C.byref(k))
        return out[: self.N * k.value].reshape(self.N, k.value)

    def terms(self):
        k = np.zeros(self.n_terms, dtype=np.int64)
        g = np.zeros(self.n_terms, dtype=np.int64)
        libs()[1].rbmhost_terms(self.h, _p(k, I64P), None, None, _p(g, I64P))
        sites = np.zeros(int(k.sum()), dtype=np.int64)
        mats = np.zeros(int(np.sum(4 ** k)) * 2)
        libs()[1].rbmhost_terms(self.h, None, _p(sites, I64P), _p(mats, DP), None)
        out, so, mo = [], 0, 0
        mz = mats.view(np.complex128)
        for t in range(self.n_terms):
            d = 1 << int(k[t])
            out.append((sites[so: so + k[t]], mz[mo: mo + d * d].reshape(d, d), int(g[t])))
            so += k[t]
            mo += d * d
        return out

    def exact_energy(self):
        return libs()[1].rbmhost_exact_energy(self.h)

    def use_evaluation(self, hexagons=True):
        libs()[1].rbmhost_use_evaluation(self.h, C.c_int32(int(hexagons)))
        self._refresh()

    def set_n_samples(self, n):
        libs()[1].rbmhost_set_n_samples(self.h, C.c_int64(n))
        self._refresh()

    def initial_weights(self, seed):
        p = np.zeros(self.P, dtype=np.complex128)
        libs()[1].rbmhost_initial_weights(self.h, C.c_uint64(seed), _p(p, DP))
        return p

    def gpu(self, precision=0, lanes=0):
        return Gpu(self, precision, lanes)


class Gpu:
    """One device context (rbmgpu_*)."""

    def __init__(self, model, precision=0, lanes=0):
        g, h = libs()
        ctx = C.c_void_p()
        rc = h.rbmhost_gpu_create(model.h, C.c_int32(precision), C.c_int32(lanes), C.byref(ctx))
        self._c = ctx.value
        if rc != 0:
            msg = g.rbmgpu_last_error(C.c_void_p(self._c)).decode() if self._c else "rbmgpu_create failed"
            if self._c:
                g.rbmgpu_destroy(C.c_void_p(self._c))
            self._c = None
            raise RuntimeError("rbmgpu: " + msg)
        self.model = model
        n = np.zeros(2, dtype=np.int64)
        g.rbmgpu_n_params(self.c, _p(n[:1], I64P), _p(n[1:], I64P))
        self.P, self.n_neural = int(n[0]), int(n[1])
        self.N = model.N
        self.n_words = model.n_words
        self.n_groups = model.n_groups
        self.n_chains = model.n_chains
        self.Hn = model.learn_hidden

    def __del__(self):
        if getattr(self, "_c", None):
            libs()[0].rbmgpu_destroy(C.c_void_p(self._c))
            self._c = None

    @property
    def c(self):
        return C.c_void_p(self._c)

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("rbmgpu: " + libs()[0].rbmgpu_last_error(self.c).decode())

    def launch_shape(self):
        """(lanes per chain, units per lane, block size)"""
        return self._layout()[:3]

    def structured(self):
        """True if S.x runs the structured path (no O matrix)"""
        return bool(self._layout()[3] & 1)

    def chains(self):
        """chains the sampler runs (the autotuned count with gpu.auto_chains)"""
        n = C.c_int64()
        self._chk(libs()[0].rbmgpu_n_chains(self.c, C.byref(n)))
        return n.value

    def wave_sampler(self):
        """True if the sampler runs one chain per wavefront (k_sample_wave)"""
        return bool(self._layout()[3] & 2)

    def _layout(self):
        o = np.zeros(4, dtype=np.int32)
        self._chk(libs()[0].rbmgpu_launch_shape(self.c, _p(o, I32P)))
        return tuple(int(x) for x in o)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.complex128)
        assert p.size == self.P
        self._chk(libs()[0].rbmgpu_set_params(self.c, _p(p, DP)))

    def get_params(self):
        p = np.zeros(self.P, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_get_params(self.c, _p(p, DP)))
        return p

    def update_weights(self, dw):
        dw = np.ascontiguousarray(dw, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_update_weights(self.c, _p(dw, DP)))

    def sample_epoch(self, epoch):
        self._chk(libs()[0].rbmgpu_sample_epoch(self.c, C.c_uint32(epoch)))

    def full_epoch(self, parallel_bits=2):
        """exact enumeration of all 2^N states weighted by |psi|^2 (rbmgpu_full_epoch)"""
        self._chk(libs()[0].rbmgpu_full_epoch(self.c, C.c_int32(parallel_bits)))

    def full_weights(self):
        S = self.n_samples()
        lp = np.zeros(S, dtype=np.complex128)
        w = np.zeros(S)
        self._chk(libs()[0].rbmgpu_full_weights(self.c, _p(lp, DP), _p(w, DP)))
        return lp, w

    def load_samples(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint64).reshape(-1, self.n_words)
        self._chk(libs()[0].rbmgpu_load_samples(self.c, _p(s, U64P), C.c_int64(s.shape[0])))

    def n_samples(self):
        n = C.c_int64()
        self._chk(libs()[0].rbmgpu_n_samples(self.c, C.byref(n)))
        return n.value

    def reduce(self, want_o=True):
        m = Moments()
        P = self.P
        om = np.zeros(P, dtype=np.complex128) if want_o else None
        ov = np.zeros(P) if want_o else None
        eo = np.zeros(P, dtype=np.complex128) if want_o else None
        gm = np.zeros(self.n_groups, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_reduce(self.c, C.byref(m), _p(om, DP), _p(ov, DP), _p(eo, DP),
                                          _p(gm, DP)))
        res = dict(mean=complex(m.e_mean[0], m.e_mean[1]), var=m.e_var, var_binned=m.e_var_binned,
                   tau=m.e_tau, std=m.e_std, n_bins=m.n_bins, acc=m.acceptance, group_means=gm)
        if want_o:
            res.update(o_mean=om, o_var=ov, eo=eo)
        return res

    def apply_S(self, x, r1, r2, rd):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.P, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_apply_S(self.c, _p(x, DP), _p(y, DP), C.c_double(r1), C.c_double(r2),
                                           C.c_double(rd)))
        return y

    def apply_S_noalloc(self, x_dev, y_dev, r1, r2, rd):
        """rbmgpu_apply_S_noalloc on DEVICE addresses x_dev, y_dev (ints, P
        complex doubles each); asynchronous on the context's stream."""
        self._chk(libs()[0].rbmgpu_apply_S_noalloc(self.c, C.c_void_p(x_dev), C.c_void_p(y_dev), C.c_double(r1),
                                                   C.c_double(r2), C.c_double(rd)))

    def train_epoch(self, epoch, lr, reg1, reg2, reg1_delta, max_iter=0, rtol=0.0, graph=True):
        """one SR epoch on the device (rbmgpu_train_epoch); returns its statistics"""
        p = SrParams(lr, reg1, reg2, reg1_delta, max_iter, rtol, int(graph))
        st = EpochStats()
        self._chk(libs()[0].rbmgpu_train_epoch(self.c, C.c_uint32(epoch), C.byref(p), C.byref(st)))
        return dict(mean=complex(st.e_mean[0], st.e_mean[1]), var=st.e_var, var_binned=st.e_var_binned,
                    tau=st.e_tau, std=st.e_std, n_bins=st.n_bins, acc=st.acceptance, itn=st.cg_itn, rs=st.cg_rs,
                    max_dw=st.max_dw, updated=bool(st.updated), graph=bool(st.graph_replayed), ms=st.ms,
                    explicit_s=bool(st.explicit_s))

    def sr_solve_ex(self, method, b, x0, r1, r2, rd, max_iter=0, rtol=0.0):
        """rbmgpu_sr_solve_ex: method 0 matrix-free CG, 1 CG on the explicit S
        (cg-single), 2 direct"""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.array(x0, dtype=np.complex128)
        itn = C.c_int64()
        rs = C.c_double()
        self._chk(libs()[0].rbmgpu_sr_solve_ex(self.c, C.c_int32(method), _p(b, DP), _p(x, DP), C.c_double(r1),
                                               C.c_double(r2), C.c_double(rd), C.c_int64(max_iter),
                                               C.c_double(rtol), C.byref(itn), C.byref(rs)))
        return x, itn.value, rs.value

    def build_S(self, r1, r2, rd):
        """explicit S with the CG regularisation (P x P)"""
        S = np.zeros(self.P * self.P, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_build_S(self.c, C.c_double(r1), C.c_double(r2), C.c_double(rd), _p(S, DP)))
        return S.reshape(self.P, self.P, order="F")

    def sr_solve(self, b, x0, r1, r2, rd, max_iter=0, rtol=0.0):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.array(x0, dtype=np.complex128)
        itn = C.c_int64()
        rs = C.c_double()
        self._chk(libs()[0].rbmgpu_sr_solve(self.c, _p(b, DP), _p(x, DP), C.c_double(r1), C.c_double(r2),
                                            C.c_double(rd), C.c_int64(max_iter), C.c_double(rtol),
                                            C.byref(itn), C.byref(rs)))
        return x, itn.value, rs.value

    def eval_logpsi(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint64).reshape(-1, self.n_words)
        out = np.zeros(s.shape[0], dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_eval_logpsi(self.c, _p(s, U64P), C.c_int64(s.shape[0]), _p(out, DP)))
        return out

    def psi_ratio(self, state, flips):
        s = np.ascontiguousarray(state, dtype=np.uint64).reshape(-1)
        f = np.ascontiguousarray(flips, dtype=np.int32)
        out = np.zeros(1, dtype=np.complex128)
        self._chk(libs()[0].rbmgpu_psi_ratio(self.c, _p(s, U64P), _p(f, I32P), C.c_int32(f.size),
                                             _p(out, DP)))
        return out[0]

    def get_samples(self, states=True, obs=True, O=False, chain_acc=False):
        S = self.n_samples()
        st = np.zeros((S, self.n_words), dtype=np.uint64) if states else None
        ob = np.zeros((S, self.n_groups), dtype=np.complex128) if obs else None
        Om = np.zeros(S * self.P, dtype=np.complex128) if O else None
        ca = np.zeros(self.chains()) if chain_acc else None
        self._chk(libs()[0].rbmgpu_get_samples(self.c, _p(st, U64P), _p(ob, DP), _p(Om, DP), _p(ca, DP)))
        if Om is not None:
            Om = Om.reshape(S, self.P).T
        return st, ob, Om, ca

    def kernel_times(self):
        """ms: sample, assemble, reduce; cg_iter = one CG iteration (S.x and
        the CG step) of the last sr_solve; sx_main = the S.x main kernel
        alone, from the last time_sx."""
        t = np.zeros(6)
        self._chk(libs()[0].rbmgpu_kernel_times(self.c, _p(t, DP)))
        return dict(zip(["sample", "assemble", "reduce", "cg_iter", "unused", "sx_main"], t.tolist()))

    # ---- learned proposal
    def learn_weights(self, params=None):
        """get (params None) or set the FP32 master weights [W1|b1|W2|b2]"""
        n = self.Hn * self.N * 2 + self.Hn + self.N
        if params is None:
            out = np.zeros(n)
            self._chk(libs()[0].rbmgpu_learn_weights(self.c, C.c_int32(0), _p(out, DP)))
            return out
        p = np.ascontiguousarray(params, dtype=np.float64)
        assert p.size == n
        self._chk(libs()[0].rbmgpu_learn_weights(self.c, C.c_int32(1), _p(p, DP)))

    def learn_grad(self, states, seeds, w, beta):
        st = np.ascontiguousarray(states, dtype=np.uint64)
        sd = np.ascontiguousarray(seeds, dtype=np.int32)
        ww = np.ascontiguousarray(w, dtype=np.float64)
        g = np.zeros(self.Hn * self.N * 2 + self.Hn + self.N)
        loss = C.c_double()
        self._chk(libs()[0].rbmgpu_learn_grad(self.c, _p(st, U64P), _p(sd, I32P), _p(ww, DP),
                                              C.c_int64(len(sd)), C.c_double(beta), _p(g, DP), C.byref(loss)))
        return loss.value, g

    def learn_logits(self, states):
        st = np.ascontiguousarray(states, dtype=np.uint64)
        n = st.shape[0]
        z = np.zeros(n * self.N)
        self._chk(libs()[0].rbmgpu_learn_logits(self.c, _p(st, U64P), C.c_int64(n), _p(z, DP)))
        return z.reshape(n, self.N)

    def learn_train(self, lr=1e-3, beta=0.01, batch=16384):
        out = np.zeros(4)
        self._chk(libs()[0].rbmgpu_learn_train(self.c, C.c_double(lr), C.c_double(beta), C.c_int64(batch),
                                               _p(out, DP)))
        return dict(loss=out[0], reward=out[1], proposed=out[2], accepted=out[3])

    def learn_set_fraction(self, f):
        self._chk(libs()[0].rbmgpu_learn_set_fraction(self.c, C.c_double(f)))

    def time_sx(self, reps=20):
        """mean ms of the main S.x kernel over reps back-to-back launches"""
        ms = C.c_double()
        self._chk(libs()[0].rbmgpu_time_sx(self.c, C.c_int32(reps), C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._chk(libs()[0].rbmgpu_synchronize(self.c))
