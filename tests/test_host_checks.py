"""Host-side checks of the Python mirror: the C ABI carries no array lengths,
so every array a spec hands to it is checked before the call, with the
reference's messages where it has one (cpp/core/signals.cpp:452, 493-497, 598,
604, 612). No device is needed."""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa


def spec(nx=6, ny=5, S=2, **kw):
    g = msa.Grid(nx, ny, 1, 1.0, 1.0, 1.0, np.zeros(nx * ny, np.uint8))
    return msa.SignalGridSpec(g, np.full(S, 0.1, np.float32), **kw)


def good_rates(nx=6, ny=5, S=2):
    return np.zeros(S * nx * ny, np.float32)


def test_a_complete_spec_passes():
    s = spec(boundary_values=np.zeros(12, np.float32), loss_rates=good_rates(),
             source_rates=good_rates(), phi=np.zeros(30, np.float32), monod_vmax=1.0,
             monod_k=0.1,
             velocity=msa.VelocityField(np.zeros(7 * 5), np.zeros(6 * 6), np.zeros(60)))
    c = s._c([])
    assert c.signal_count == 2


@pytest.mark.parametrize("kw", [dict(loss_rates=good_rates()), dict(source_rates=good_rates()),
                                dict(loss_rates=good_rates()[:-1], source_rates=good_rates()),
                                dict(loss_rates=good_rates(), source_rates=good_rates()[1:])])
def test_affine_reaction_needs_both_arrays_of_the_level_count(kw):
    with pytest.raises(msa.InvalidArgument,
                       match="affine reaction arrays must match the grid level count"):
        spec(**kw)._c([])


def test_boundary_values_must_match_the_signal_count():
    with pytest.raises(msa.InvalidArgument, match="must match signal count"):
        spec(boundary_values=np.zeros(6, np.float32))._c([])
    with pytest.raises(msa.InvalidArgument, match="must match signal count"):
        spec(patches=((0, (0, 0), (2, 1), [1.0]),))._c([])


@pytest.mark.parametrize("which", [0, 1, 2])
def test_velocity_field_must_cover_every_face(which):
    faces = [np.zeros(7 * 5), np.zeros(6 * 6), np.zeros(60)]
    faces[which] = faces[which][:-1]
    with pytest.raises(msa.InvalidArgument, match="must cover every lattice face"):
        spec(velocity=msa.VelocityField(*faces))._c([])


def test_occupancy_and_monod_constants_are_checked():
    with pytest.raises(msa.InvalidArgument, match="one value per grid site"):
        spec(phi=np.zeros(29, np.float32), monod_vmax=1.0, monod_k=0.1)._c([])
    with pytest.raises(msa.InvalidArgument, match="one value per signal"):
        spec(phi=np.zeros(30, np.float32), monod_vmax=[1.0, 2.0, 3.0], monod_k=0.1)._c([])


def test_obstacle_mask_must_cover_every_site():
    g = msa.Grid(6, 5, 1, 1.0, 1.0, 1.0, np.zeros(29, np.uint8))
    with pytest.raises(msa.InvalidArgument, match="obstacle mask must cover every site"):
        g._c([])
