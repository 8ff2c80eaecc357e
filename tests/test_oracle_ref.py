"""CPU: the Green-tensor face integrals of the oracle (whose rules the device
assembly evaluates) against the reference's own C++ cores, compiled from the
reference sources into oracle/_ref by oracle/build_ref.sh: the DEMCEM
singular square-pair values and the reference's 4/10/20-point cubature
(assembly_nop_const.m:16-168, cubature_nop_const.cpp:8-125,
direct_ws_{st,ea,va}_const.cpp)."""
import numpy as np
import pytest

from oracle import greens as ogr, ref_greens as rg

pytestmark = pytest.mark.skipif(not rg.available(), reason='oracle/_ref/libvpref.so not built')

DX = 20e-9                                     # the straight-waveguide configuration
KO = 2 * np.pi * np.sqrt(2.085) / 1550e-9      # wavenumber in SiO2 at 1550 nm


def test_contact_classifier_reproduces_the_correction_tables():
    # nearby_correct.m's C tables for m = [2,1,1] and [2,1,2] (offsets (1,0,0)
    # and (1,0,1)); rows: observation face of cell m, columns: source face of cell 0
    tables = {
        (1, 0, 0): [[None, 'ST', 'EAo', 'EAo', 'EAo', 'EAo'],
                    [None] * 6,
                    [None, 'EAo', 'EAc', None, 'VAo', 'VAo'],
                    [None, 'EAo', None, 'EAc', 'VAo', 'VAo'],
                    [None, 'EAo', 'VAo', 'VAo', 'EAc', None],
                    [None, 'EAo', 'VAo', 'VAo', None, 'EAc']],
        (1, 0, 1): [[None, 'EAc', 'VAo', 'VAo', None, 'EAo'],
                    [None] * 6,
                    [None, 'VAo', 'VAc', None, None, 'VAo'],
                    [None, 'VAo', None, 'VAc', None, 'VAo'],
                    [None, 'EAo', 'VAo', 'VAo', None, 'EAc'],
                    [None] * 6],
    }
    for m, tab in tables.items():
        for p in range(6):
            for q in range(6):
                assert rg.contact(m, p, q) == tab[p][q], (m, p, q)
    # the self cell: identical faces, opposite faces apart, the rest share an edge
    for p in range(6):
        for q in range(6):
            expect = 'ST' if p == q else (None if p // 2 == q // 2 else 'EAo')
            assert rg.contact((0, 0, 0), p, q) == expect


def test_singular_pairs_match_reference_demcem():
    kappa = KO * DX
    S = rg.singular_values(KO, DX)
    seen = set()
    for m in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
        I = ogr.face_pair_integrals(m, kappa)
        for p in range(6):
            for q in range(6):
                t = rg.contact(m, p, q)
                if t is None:
                    continue
                seen.add(t)
                ref = S[t] / DX ** 3
                assert abs(I[p, q] - ref) <= 1e-12 * abs(ref), (m, p, q, t, I[p, q], ref)
    assert seen == {'ST', 'EAc', 'EAo', 'VAc', 'VAo'}


@pytest.mark.parametrize('m', [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_touching_cells_match_reference_rules(m):
    # regular pairs by the reference's 20-point rule, singular ones by DEMCEM
    Jr = rg.J_ref(m, KO, DX) / DX ** 3
    Jo = ogr.face_pair_integrals(m, KO * DX)
    assert np.abs(Jr - Jo).max() <= 1e-12 * np.abs(Jo).max()


def test_green_tensor_against_reference_rules():
    # G_mn on a 7 x 3 x 2 block of offsets: touching cells agree to rounding;
    # further cells to the accuracy of the reference's 10- and 4-point rules
    L, M, N = 7, 3, 2
    G = ogr.assemble_G(L, M, N, DX, KO)
    offs = [(i, j, k) for k in range(N) for j in range(M) for i in range(L)]
    Gr = rg.G_ref(offs, KO, DX)
    scale = np.abs(G).max()
    for (i, j, k), g in zip(offs, Gr):
        d = np.abs(G[i, j, k] - g).max() / scale
        if max(i, j, k) <= 1:
            assert d <= 1e-12, ((i, j, k), d)
        else:
            assert d <= 1e-6, ((i, j, k), d)
