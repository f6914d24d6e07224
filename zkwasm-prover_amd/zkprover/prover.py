"""Host binding of the prover core (zkp_pk_* / zkp_job_* / zkp_prove).

A `Circuit` describes the PLONKish benchmark circuit; `ProvingKey` holds its
resident shape data on the device; `Prover` owns one in-flight proof's device
state and runs proofs on a stream.  The arithmetic is native (prover_core.cpp
over the kernel library); this module only marshals arguments.
"""
import ctypes

import numpy as np

from ._lib import HipError, check, lib


class Circuit(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("k", "advice", "fixed", "instance", "lookups", "table_log",
                                             "perm_cols", "shuffles", "shuffle_log", "blinding")]

    @classmethod
    def benchmark(cls, k):
        """the 40-advice circuit family of the k=20 benchmark configuration"""
        return cls(k=k, advice=40, fixed=10, instance=1, lookups=6, table_log=16, perm_cols=12, shuffles=2,
                   shuffle_log=min(18, k - 2), blinding=5)


class ProofInfo(ctypes.Structure):
    _fields_ = [("ncommit", ctypes.c_int), ("z_ok", ctypes.c_int), ("quotient_ok", ctypes.c_int),
                ("z_fail", ctypes.c_int), ("open_ok", ctypes.c_int), ("nevals", ctypes.c_int),
                ("phase_ms", ctypes.c_double * 7), ("total_ms", ctypes.c_double),
                ("transcript", ctypes.c_ubyte * 32), ("challenges", (ctypes.c_ulonglong * 4) * 4)]

    PHASES = ("witness", "advice_commit_intt", "multiplicities", "grand_products", "h_eval", "vanishing",
              "multiopen")

    def as_dict(self):
        return {"ncommit": self.ncommit, "z_ok": bool(self.z_ok), "quotient_ok": bool(self.quotient_ok),
                "open_ok": bool(self.open_ok), "nevals": self.nevals,
                "phase_ms": {k: round(self.phase_ms[i], 3) for i, k in enumerate(self.PHASES)},
                "total_ms": round(self.total_ms, 3)}


BATCHED = 1
TAMPER_GATE = 2
SHPLONK = 4
NO_OPEN = 8
TAMPER_EVAL = 16
HOST_WITNESS = 32
GRAPH = 64


def _decl():
    L = lib()
    L.zkp_pk_bytes.restype = ctypes.c_size_t
    L.zkp_pk_bytes.argtypes = [ctypes.c_void_p]
    L.zkp_job_bytes.restype = ctypes.c_size_t
    L.zkp_job_bytes.argtypes = [ctypes.c_void_p]
    L.zkp_commitment_count.restype = ctypes.c_int
    L.zkp_commitment_count.argtypes = [ctypes.c_void_p]
    L.zkp_pk_create.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_void_p]
    L.zkp_pk_destroy.argtypes = [ctypes.c_void_p]
    L.zkp_job_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.zkp_job_destroy.argtypes = [ctypes.c_void_p]
    L.zkp_prove.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p]
    L.zkp_pk_srs_secret.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.zkp_pk_fixed_commitments.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.zkp_pk_fixed_commitments.restype = ctypes.c_int
    L.zkp_proof_openings.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.zkp_proof_openings.restype = ctypes.c_int
    L.zkp_proof_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    L.zkp_proof_bytes.restype = ctypes.c_longlong
    L.zkp_proof_instance.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.zkp_pk_vk_repr.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.zkp_set_host_witness.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.zkp_proof_witness.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return L


def footprint(circuit):
    """(key bytes, per-proof bytes) of device memory"""
    L = _decl()
    return L.zkp_pk_bytes(ctypes.byref(circuit)), L.zkp_job_bytes(ctypes.byref(circuit))


class ProvingKey:
    def __init__(self, circuit, seed=1, stream=0):
        self.L = _decl()
        self.circuit = circuit
        self.handle = ctypes.c_void_p()
        check(self.L.zkp_pk_create(ctypes.byref(circuit), seed, stream, ctypes.byref(self.handle)), "zkp_pk_create")
        self.ncommit = self.L.zkp_commitment_count(ctypes.byref(circuit))

    def srs_secret(self):
        """the synthetic setup's tau (Montgomery Fr, (1, 4) u64): test use only"""
        out = np.zeros((1, 4), dtype=np.uint64)
        check(self.L.zkp_pk_srs_secret(self.handle, out.ctypes.data), "zkp_pk_srs_secret")
        return out

    def vk_repr(self):
        """the verifying key's transcript representation (Montgomery Fr, (1, 4) u64)"""
        out = np.zeros((1, 4), dtype=np.uint64)
        check(self.L.zkp_pk_vk_repr(self.handle, out.ctypes.data), "zkp_pk_vk_repr")
        return out

    def fixed_commitments(self):
        """verifying-key commitments of the fixed columns (q, t, f_j) and the
        sigma columns, (fixed + perm_cols, 8) affine"""
        out = np.zeros((self.circuit.fixed + self.circuit.perm_cols, 8), dtype=np.uint64)
        cnt = self.L.zkp_pk_fixed_commitments(self.handle, out.ctypes.data, len(out))
        if cnt < 0:
            raise HipError("zkp_pk_fixed_commitments")
        return out[:cnt]

    def close(self):
        if self.handle:
            check(self.L.zkp_pk_destroy(self.handle), "zkp_pk_destroy")
            self.handle = ctypes.c_void_p()


class Prover:
    def __init__(self, pk):
        self.pk = pk
        self.L = pk.L
        self.handle = ctypes.c_void_p()
        check(self.L.zkp_job_create(pk.handle, ctypes.byref(self.handle)), "zkp_job_create")

    def prove(self, seed, flags=BATCHED, stream=0, want_pieces=True, pieces_out=None):
        """returns (commitments (ncommit, 8) u64 affine in transcript order -- instance,
        advice, m, permutation z, logup z, shuffle z, random, quotient pieces, opening --
        h_pieces (n, 4) or None, ProofInfo);
        pieces_out: a caller's (n, 4) u64 array to receive h_pieces (e.g. pinned memory)"""
        n = 1 << self.pk.circuit.k
        com = np.zeros((self.pk.ncommit, 8), dtype=np.uint64)
        if pieces_out is not None:
            if pieces_out.shape != (n, 4) or pieces_out.dtype != np.uint64 or not pieces_out.flags.c_contiguous:
                raise ValueError("pieces_out must be a contiguous (%d, 4) uint64 array" % n)
            pieces, want_pieces = pieces_out, True
        else:
            pieces = np.zeros((n, 4), dtype=np.uint64) if want_pieces else None
        info = ProofInfo()
        check(self.L.zkp_prove(self.handle, seed, flags, com.ctypes.data,
                               pieces.ctypes.data if want_pieces else None, ctypes.byref(info), stream),
              "zkp_prove")
        return com, pieces, info

    def proof_bytes(self):
        """the last proof's bytes (written points compressed, written scalars)"""
        ln = self.L.zkp_proof_bytes(self.handle, None, 0)
        if ln < 0:
            raise HipError("zkp_proof_bytes")
        out = np.zeros(max(ln, 1), dtype=np.uint8)
        if self.L.zkp_proof_bytes(self.handle, out.ctypes.data, len(out)) != ln:
            raise HipError("zkp_proof_bytes")
        return out[:ln].tobytes()

    def instance(self):
        """the last proof's instance columns, (instance, n, 4) Montgomery"""
        n = 1 << self.pk.circuit.k
        out = np.zeros((max(self.pk.circuit.instance, 1), n, 4), dtype=np.uint64)
        check(self.L.zkp_proof_instance(self.handle, out.ctypes.data, n), "zkp_proof_instance")
        return out[:self.pk.circuit.instance]

    def set_host_witness(self, advice_ptrs, instance_ptrs):
        """host columns (2^k Fr each) the next HOST_WITNESS proof uploads: lists
        of addresses, kept by the library until that proof returns"""
        adv = (ctypes.c_void_p * max(len(advice_ptrs), 1))(*advice_ptrs)
        ins = (ctypes.c_void_p * max(len(instance_ptrs), 1))(*instance_ptrs)
        self._host_witness = (adv, ins)
        check(self.L.zkp_set_host_witness(self.handle, adv, len(advice_ptrs), ins, len(instance_ptrs)),
              "zkp_set_host_witness")

    def witness_to_host(self, advice_ptrs, instance_ptrs):
        """copies the last proof's advice and instance columns to host addresses"""
        adv = (ctypes.c_void_p * max(len(advice_ptrs), 1))(*advice_ptrs)
        ins = (ctypes.c_void_p * max(len(instance_ptrs), 1))(*instance_ptrs)
        check(self.L.zkp_proof_witness(self.handle, adv, ins), "zkp_proof_witness")

    def openings(self):
        """the last proof's evaluations ((count, 4) Montgomery; h(x) first, then
        the written ones in the reference's order) and its opening challenges
        [v, y', u] ((3, 4))"""
        cnt = self.L.zkp_proof_openings(self.handle, None, 0, None)
        ev = np.zeros((max(cnt, 1), 4), dtype=np.uint64)
        ch = np.zeros((3, 4), dtype=np.uint64)
        if self.L.zkp_proof_openings(self.handle, ev.ctypes.data, len(ev), ch.ctypes.data) != cnt:
            raise HipError("zkp_proof_openings")
        return ev[:cnt], ch

    def close(self):
        if self.handle:
            check(self.L.zkp_job_destroy(self.handle), "zkp_job_destroy")
            self.handle = ctypes.c_void_p()


__all__ = ["Circuit", "ProofInfo", "ProvingKey", "Prover", "footprint", "BATCHED", "TAMPER_GATE", "SHPLONK",
           "NO_OPEN", "TAMPER_EVAL", "HOST_WITNESS", "GRAPH", "HipError"]
