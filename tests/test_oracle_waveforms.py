"""Pins of the oracle's source waveforms and Yee-position voxelization by closed forms.

The broadband ('wavelength' time definition) branch of the oracle's pulse_parameters and
source_time_signal drives every configuration's source.  Its checks here are the invariants the
reference's own tests hold for that branch (tests/test_broadband.py:18-56): the carrier at the
band midpoint, the 1/4 short-pulse width ratio, the chirp threshold, an instantaneous frequency
that sweeps monotonically from the low to the high band edge, flat at +-2 sigma, and a tapered
envelope that is C1 with compact support.  The instantaneous frequency is measured from the
analytic signal (sine and cosine phases), not from the formula, so a dropped term, a wrong sign
of the chirp or a swapped band edge fails.

The 'yee' sampling is pinned by a rectangle whose edges fall between the half-cell and nodal Yee
positions, so each E component owns a different, hand-counted box of cells.
"""
import math

import numpy as np
import pytest

from oracle.scene import (C0, SOURCE_DEFAULT, TAIL_INNER, TAIL_OUTER, Scene, gaussian_envelope, pulse_parameters,
                          source_time_signal)


def broad(**kw):
    s = dict(SOURCE_DEFAULT)
    s.update(pulse='broadband', time_definition='wavelength', wavelength_start=1.2, wavelength_stop=1.8)
    s.update(kw)
    return s


def test_broadband_carrier_width_and_chirp_threshold():
    s = broad(optimize_for_short_pulse=False)
    f, sigma, offset, span, chirped = pulse_parameters(s)
    assert f == pytest.approx(.5*(C0/1.2e-6+C0/1.8e-6), rel=1e-15)
    assert span == pytest.approx(C0/1.2e-6-C0/1.8e-6, rel=1e-15)
    assert sigma > 0 and offset > 3*sigma and chirped
    # Full width at half maximum of the intensity: exp(-u^2) = 1/2 at u = length/(2 sigma) ...
    # the reference's length 8/(f_high + 0.01 f_low), sigma = length/(2 sqrt(ln 2)).
    length = 8/(C0/1.2e-6+.01*C0/1.8e-6)
    assert math.exp(-(length/2/sigma)**2) == pytest.approx(.5, rel=1e-12)
    assert pulse_parameters(broad(optimize_for_short_pulse=True))[1] == sigma/4
    # Narrow band (1.4-1.6 um) stays below the chirp threshold; a single wavelength has no span.
    assert not pulse_parameters(broad(optimize_for_short_pulse=True, wavelength_start=1.4, wavelength_stop=1.6))[4]
    *_, span0, chirped0 = pulse_parameters(broad(wavelength_start=1.55, wavelength_stop=1.55))
    assert span0 == 0 and not chirped0


def instantaneous_frequency(s, t):
    sine = source_time_signal(s, t)
    cosine = source_time_signal(dict(s, phase=90.), t)
    phase = np.unwrap(np.arctan2(sine, cosine))
    return -np.gradient(phase, t)/(2*np.pi)


def test_chirp_sweeps_between_the_band_edges():
    s = broad(optimize_for_short_pulse=False)
    f, sigma, offset, span, _ = pulse_parameters(s)
    t = offset+np.linspace(-3, 3, 6001)*sigma
    freq = instantaneous_frequency(s, t)
    assert freq[500] == pytest.approx(C0/1.8e-6, rel=1e-9)       # u = -2.5: low edge
    assert freq[5500] == pytest.approx(C0/1.2e-6, rel=1e-9)      # u = +2.5: high edge
    assert freq[3000] == pytest.approx(f, rel=1e-9)              # u = 0: midpoint
    assert np.min(np.diff(freq)) > -1e-11*f                       # monotone sweep
    for center in (-2, 2):                                        # C1 join of the cosine sweep and the linear tails
        k = int(np.argmin(abs((t-offset)/sigma-center)))
        assert abs(freq[k+1]-freq[k-1]) < span*1e-6


def test_unchirped_broadband_and_cycles_pulses_have_a_constant_carrier():
    s = broad(optimize_for_short_pulse=True, wavelength_start=1.5, wavelength_stop=1.6)
    f, sigma, offset, _, chirped = pulse_parameters(s)
    assert not chirped
    t = offset+np.linspace(-2, 2, 2001)*sigma
    assert np.allclose(instantaneous_frequency(s, t)[5:-5], f, rtol=1e-8)
    c = dict(SOURCE_DEFAULT, wavelength=1.55, pulse_cycles=3.)
    f, sigma, offset, _, chirped = pulse_parameters(c)
    assert f == C0/1.55e-6 and sigma == 3/f and offset == 4*sigma and not chirped
    t = offset+np.linspace(-2, 2, 2001)*sigma
    # The cycles definition runs its carrier from t = 0: sin(w t); frequency -w/(2 pi) under the
    # broadband sign convention of instantaneous_frequency.
    assert np.allclose(-instantaneous_frequency(c, t)[5:-5], f, rtol=1e-8)
    np.testing.assert_allclose(source_time_signal(c, t), np.exp(-.5*((t-offset)/sigma)**2)*np.sin(2*np.pi*f*t),
                               rtol=0, atol=1e-15)


def test_tapered_envelope_is_c1_with_compact_support():
    h = 1e-5
    for boundary in (TAIL_INNER, TAIL_OUTER):
        x = np.array([boundary-h, boundary, boundary+h])
        slopes = np.diff(gaussian_envelope(x, True))/h
        assert abs(slopes[1]-slopes[0]) < 3e-8
    assert gaussian_envelope(np.array([0, TAIL_OUTER, TAIL_OUTER+1]), True).tolist() == [1, 0, 0]
    assert gaussian_envelope(np.array([TAIL_INNER]), True)[0] == pytest.approx(1e-4)
    assert gaussian_envelope(np.array([-3.]), False)[0] == math.exp(-4.5)


def test_yee_sampling_of_a_rectangle_is_the_closed_form():
    """Region 2.0 x 1.6 x 1.2 um at mesh 0.1 (20 x 16 x 12 cells); rectangle of 0.52 x 0.32 x 0.22 um
    at the centre.  Half-cell positions inside: x 6 (+-0.05, 0.15, 0.25), y 4, z 2; nodal positions
    inside: x 5 (0, +-0.1, 0.2), y 3, z 3.  Ex sits half a cell along x, Ey along y, Ez along z."""
    project = {'region': {'dimension': '3d', 'size': (2.0, 1.6, 1.2), 'mesh': .1, 'material_sampling': 'yee',
                          'background_index': 1.0},
               'structures': [{'kind': 'rectangle', 'size': (.52, .32, .22), 'material': 'Si (constant n)'}]}
    eps = Scene(project).voxelize()
    assert eps.shape == (20, 16, 12, 3)
    core = np.float32(3.48**2)
    boxes = {0: (slice(7, 13), slice(7, 10), slice(5, 8)),     # Ex: half x, nodal y and z
             1: (slice(8, 13), slice(6, 10), slice(5, 8)),     # Ey: half y
             2: (slice(8, 13), slice(7, 10), slice(5, 7))}     # Ez: half z
    for comp, box in boxes.items():
        expected = np.ones((20, 16, 12), np.float32)
        expected[box] = core
        np.testing.assert_array_equal(eps[..., comp], expected, err_msg='xyz'[comp])
    assert [int((eps[..., c] == core).sum()) for c in range(3)] == [54, 60, 30]
    project['region']['material_sampling'] = 'cell'
    cell = Scene(project).voxelize()
    assert cell.shape == (20, 16, 12) and int((cell == core).sum()) == 6*4*2
    assert np.all(cell[7:13, 6:10, 5:7] == core)


def test_mesh_order_priority_and_sphere_closed_form():
    """Lower mesh_order wins, and the later object wins at equal order; a sphere holds the cell
    centres within its radius (counted by brute force over the integer lattice)."""
    project = {'region': {'dimension': '3d', 'size': (2.0, 2.0, 2.0), 'mesh': .1, 'background_index': 1.0},
               'structures': [{'kind': 'sphere', 'radius': .45, 'material': 'Si (constant n)', 'mesh_order': 2},
                              {'kind': 'rectangle', 'size': (.2, .2, .2), 'material': 'Air', 'mesh_order': 1},
                              {'kind': 'rectangle', 'size': (2.0, .2, .2), 'center': (0, .6, 0),
                               'material': 'SiN (constant n)', 'mesh_order': 2},
                              {'kind': 'rectangle', 'size': (2.0, .2, .2), 'center': (0, .6, 0),
                               'material': 'SiO2 (constant n)', 'mesh_order': 2}]}
    eps = Scene(project).voxelize()
    i = np.arange(20)-9.5                                   # centres in units of the mesh
    r2 = i[:, None, None]**2+i[None, :, None]**2+i[None, None, :]**2
    sphere = r2 <= 4.5**2
    box = np.zeros_like(sphere)
    box[9:11, 9:11, 9:11] = True
    np.testing.assert_array_equal(eps == np.float32(3.48**2), sphere & ~box)
    assert np.all(eps[9:11, 9:11, 9:11] == 1)
    assert np.all(eps[:, 15:17, 9:11] == np.float32(1.444**2))
