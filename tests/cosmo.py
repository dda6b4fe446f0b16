"""Shared cosmologies and table builders for the tests."""
import functools

import numpy as np

import boltmi
import pyoracle


@functools.lru_cache(maxsize=None)
def background(kind):
    if kind == "default":        # CosmoParams() (Bolt.jl:56-66), the cell_camb test cosmology
        par = boltmi.CosmoParams()
    elif kind == "nonu":         # CosmoParams(Σm_ν=0.0), the CLASS comparison (runtests.jl:85)
        par = boltmi.CosmoParams(Σm_ν=0.0)
    elif kind == "c1":           # the plain-method workload: n_s = 0.96, A_s = 2.1e-9, Σm_ν = 0
        par = boltmi.CosmoParams(h=0.7, Ω_b=0.046, Ω_c=0.224, n=0.96, A=2.1e-9, N_ν=3.046, Σm_ν=0.0)
    elif kind == "massive":      # one massive neutrino at 0.2 eV
        par = boltmi.CosmoParams(h=0.67, Ω_b=0.049, Ω_c=0.26, n=0.965, A=2.1e-9, Σm_ν=0.2 * boltmi.MASS_NATURAL)
    else:
        raise KeyError(kind)
    bg = boltmi.Background(par)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    return par, bg, ih


@functools.lru_cache(maxsize=None)
def massive_nq(nq):
    """the one-massive-neutrino cosmology with nq momentum bins: tables, ionization
    history and the oracle on those tables"""
    par = boltmi.CosmoParams(h=0.67, Ω_b=0.049, Ω_c=0.26, n=0.965, A=2.1e-9, Σm_ν=0.2 * boltmi.MASS_NATURAL)
    bg = boltmi.Background(par, nq=nq)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    return par, bg, ih, pyoracle.Oracle(par.as_tuple(), bg.tables_dict())


@functools.lru_cache(maxsize=None)
def oracle(kind):
    par, bg, ih = background(kind)
    return pyoracle.Oracle(par.as_tuple(), bg.tables_dict())


def ocfg(ℓᵧ=8, ℓ_ν=8, ℓ_mν=10, nq=15, reltol=1e-6, abstol=1e-6, refine=1):
    return pyoracle.Oracle.cfg(ℓᵧ, ℓ_ν, ℓ_mν, nq, reltol, abstol, 0, refine)


def xnode(j, x0=-20.0, x1=0.0, nx=2001):
    return (x0 * (nx - 1 - j) + x1 * j) / (nx - 1)
