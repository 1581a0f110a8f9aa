"""Shared builders for the parity tests (meshes, oracle matrices)."""
from __future__ import annotations

import functools
import math

import numpy as np
import scipy.sparse as sp

import oracle
from fem25_amd import configs as cf


@functools.lru_cache(maxsize=None)
def mesh_case(name: str):
    """(mesh, model, pml) for the named test case."""
    if name == "tiny":          # 6 x 4, p = 4, all interior
        m = cf.structured_mesh(6, 4, 4, width=12e3, depth=8e3)
        return m, cf.homogeneous_model(m), cf.PML()
    if name == "tiny_layered":  # 6 x 4, p = 4, heterogeneous
        m = cf.structured_mesh(6, 4, 4, width=12e3, depth=8e3)
        return m, cf.layered_model(m, seed=7, n_anom=4, amp=0.1), cf.PML()
    if name == "pml":           # 10 x 6, p = 4, 2 PML + 1 Bielak rings
        m = cf.structured_mesh(10, 6, 4, width=20e3, depth=12e3, npml=2, nbielak=1)
        return m, cf.layered_model(m, seed=3, n_anom=3), cf.PML(rcoeff=1e-3, kappa=1.0, fcent=0.05)
    if name == "p6":            # 4 x 3, p = 6
        m = cf.structured_mesh(4, 3, 6, width=8e3, depth=6e3)
        return m, cf.layered_model(m, seed=5, n_anom=2), cf.PML()
    if name == "p5":            # 5 x 3, p = 5
        m = cf.structured_mesh(5, 3, 5, width=10e3, depth=6e3)
        return m, cf.layered_model(m, seed=6, n_anom=2), cf.PML()
    if name == "p2_pml":        # 8 x 6, p = 2 with PML
        m = cf.structured_mesh(8, 6, 2, width=16e3, depth=12e3, npml=2, nbielak=1)
        return m, cf.layered_model(m, seed=9, n_anom=2), cf.PML(rcoeff=1e-3, fcent=0.1)
    if name == "c1":
        m = cf.structured_mesh(40, 20, 4)
        return m, cf.homogeneous_model(m), cf.PML()
    raise KeyError(name)


@functools.lru_cache(maxsize=None)
def oracle_graph(name: str):
    m, _, _ = mesh_case(name)
    return oracle.graph(m)


def csr(g, vals) -> sp.csr_matrix:
    return sp.csr_matrix((vals, g.jcptr - 1, g.irptr - 1), shape=(g.ndof, g.ndof))


def oracle_S(name: str, freq: float, py: float) -> sp.csr_matrix:
    m, mod, pml = mesh_case(name)
    g = oracle_graph(name)
    return csr(g, oracle.asmble_z(m, mod, g, freq, py, pml))


def relerr(a, b) -> float:
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def pairs(nfreq=2, nky=3, fmin=0.02, fmax=0.1, vs_min=3460.0):
    om, ky, f, py = cf.freq_ky_batch(nfreq, nky, fmin, fmax, vs_min)
    return om, ky, f, py


TWOPI = 2 * math.pi


def ps_inv(one):
    """GRADPTRS's PS structure from the single-variable one: each column
    copied forward for the second variable, MYGRAD = 2 (INPINV - 1) + IVINV
    (gengmask.f90:1078-1098), without the NNPG*NCON*NEE bound on ICSC_FDIST
    that stops the reference once about 60% of the anchors are inverted."""
    cols = [one.icsc[one.jcsc[c] - 1:one.jcsc[c + 1] - 1] for c in range(one.nnpgl)]
    icsc = np.concatenate([r for r in cols for _ in range(2)]).astype(np.int32)
    lens = np.repeat([len(r) for r in cols], 2)
    jcsc = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    mygrad = np.empty(2 * one.nnpgl, np.int32)
    mygrad[0::2] = 2 * one.mygrad - 1
    mygrad[1::2] = 2 * one.mygrad
    return oracle.Inv(one.nnpinv, one.ncon, 2 * one.nnpgl, int(icsc.size), one.mbufrhs, 2 * one.na35,
                      one.maskg, one.mconn, mygrad, jcsc, icsc, one.elem_wts)
