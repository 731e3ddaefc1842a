"""The exact hemisphere series of the radiation limits (tests/sphere_limits.py), pinned by its closed
forms: heave at omega = infinity and surge at omega = 0 are the translating sphere (half the
displaced mass, 0.5 of the hemisphere's m0), and the other two lie between the known bounds."""

from sphere_limits import hemisphere_added_mass


def test_translating_sphere_cases_are_one_half():
    assert abs(hemisphere_added_mass("heave", "infinity") - 0.5) < 1e-12
    assert abs(hemisphere_added_mass("surge", "zero") - 0.5) < 1e-12


def test_other_limits_converged_and_positive():
    # the rigid-lid heave (extended velocity with a kink at the waterline: fast convergence) and the
    # phi = 0 surge (a jump there: the terms decay algebraically, 2e-5 left at 240 terms)
    a = hemisphere_added_mass("heave", "zero", terms=480, points=800)
    b = hemisphere_added_mass("heave", "zero", terms=240, points=800)
    assert a > 0.0 and abs(a - b) < 1e-8 * a
    a = hemisphere_added_mass("surge", "infinity", terms=480, points=800)
    b = hemisphere_added_mass("surge", "infinity", terms=240, points=800)
    assert a > 0.0 and abs(a - b) < 1e-4 * a
