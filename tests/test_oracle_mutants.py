"""Each closed-form check of the oracle fails when the term it covers is wrong.

The oracle is rebuilt with one planted mistake at a time (ORC_MUTANT in
oracle/slab_oracle.c) and the checks that cover that term are run on it: every
one of them must fail.  The ordinary build passes the same checks
(tests/test_oracle_closed_forms.py), so together they show that the strain
trace / height update (simulation.cpp:924-937), the FLIP/PIC blend (:932) and
the depth-averaged pressure (:753, :792) are pinned by something other than a
restatement of the same formula.
"""
import pytest

import test_oracle_closed_forms as cf
from oracle import oracle as orc


def _dam_flat():
    cf.test_pressure_term_momentum_balance_of_a_dam_break(0.0, 0.01)


def _dam_tilted():
    cf.test_pressure_term_momentum_balance_of_a_dam_break(45.0, 0.03)


MUTANTS = [
    (1, "strain trace without the y term", [cf.test_height_update_follows_the_strain_trace]),
    (2, "strain trace without dt", [cf.test_height_update_follows_the_strain_trace]),
    (3, "h multiplied by (1 + tr)", [cf.test_height_update_follows_the_strain_trace]),
    (4, "FLIP/PIC weights swapped", [cf.test_flip_pic_blend_weights]),
    (5, "FLIP term dropped", [cf.test_sliding_block_coulomb_acceleration]),
    (6, "pressure without the 1/2", [_dam_flat, _dam_tilted]),
    (7, "pressure with cos instead of cos^2", [_dam_tilted]),
    (8, "pressure without h", [_dam_flat, _dam_tilted]),
    (9, "pressure with the opposite sign", [_dam_flat, _dam_tilted]),
]


@pytest.mark.parametrize("n,what,checks", MUTANTS, ids=[m[1] for m in MUTANTS])
def test_closed_form_checks_catch_planted_mistake(n, what, checks):
    path = orc.build_mutant(n)
    with orc.using_library(path):
        for check in checks:
            with pytest.raises(AssertionError):
                check()
