"""pgrav_cu -- mirror of nln/py_pgrav_cu/pgrav_cu.pyx on libpgrav.

``init()`` selects the device (CUDA_Init); ``monopole(p, accp, stride, f)``
evaluates m sinks against SoA-8 monopole records with the pM kernel
(pinteractCUDA with sz = 4, nln/py_pgrav_cu/pgrav.cu:510-564).  The reference
binding does not build as shipped (it calls an undeclared qallocCUDA and passes
8 arguments to a 7-argument declaration); this module implements the evident
intent.
"""
import pgrav
from pgrav._lib import check, lib


def init():
    lib().CUDA_Init(None)
    check()


def monopole(p, accp, stride, f):
    pgrav.mn("pM", p, accp, stride, f, layout=1)
