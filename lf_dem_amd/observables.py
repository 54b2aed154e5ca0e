"""Per-configuration observables of LF_DEM's data output, from a system's
pair list (SystemHelperFunctions.h):

* min_gap: the smallest reduced gap 2 r / (a0 + a1) - 2 over the
  interactions, starting from the interaction range (evaluateMinGap);
* potential_energy: contact springs 0.5 kn' h^2 (+ 0.5 kt' |disp_tan|^2 on
  frictional contacts), with h the overlap (minus the reduced gap),
  kn' = ((a0 + a1) / 2)^2 kn and kt' = (a0 + a1) / 2 kt
  (Contact.cpp:36-44, 469-486), plus the repulsion energy of every
  interaction, a0 a1 / (a0 + a1) * lambda * exp(-g / lambda) for a gap
  g = r - a0 - a1 > 0 and a0 a1 / (a0 + a1) * lambda * (1 - g) in contact
  (RepulsiveForce.cpp:23-40, 133-150) (getPotentialEnergy);
* contact_numbers: contacts and frictional contacts per particle, 2 n / N
  (countNumberOfContact, Simulation.cpp:722-727);
* max_disp_tan: the largest tangential displacement (evaluateMaxDispTan).

`pairs` is the dict of System.pairs() (p0, p1, state, r, disp_tan), with
state 0 for no contact, 1 frictionless, 2 and 3 frictional; `radius` the
particle radii; values in the system's internal units.
"""
import numpy as np


def _geometry(pairs, radius):
    a0 = np.asarray(radius)[pairs["p0"]]
    a1 = np.asarray(radius)[pairs["p1"]]
    return a0, a1, np.asarray(pairs["r"])


def min_gap(pairs, radius, interaction_range):
    a0, a1, r = _geometry(pairs, radius)
    if len(r) == 0:
        return float(interaction_range)
    return float(min(interaction_range, np.min(2 * r / (a0 + a1) - 2)))


def potential_energy(pairs, radius, kn, kt, repulsion, repulsive_length):
    a0, a1, r = _geometry(pairs, radius)
    state = np.asarray(pairs["state"])
    ro12 = (a0 + a1) / 2
    contact = state >= 1
    overlap = -(2 * r / (a0 + a1) - 2)
    e = np.where(contact, 0.5 * ro12 * ro12 * kn * overlap * overlap, 0.0)
    sq_tan = np.sum(np.asarray(pairs["disp_tan"]) ** 2, axis=1)
    e = e + np.where(state >= 2, 0.5 * ro12 * kt * sq_tan, 0.0)
    if repulsion > 0:
        gf = a0 * a1 / (a0 + a1)
        g = r - a0 - a1
        e = e + np.where(g > 0, gf * repulsive_length * np.exp(-np.maximum(g, 0) / repulsive_length),
                         gf * repulsive_length * (1 - g))
    return float(np.sum(e))


def contact_numbers(pairs, np_):
    state = np.asarray(pairs["state"])
    return 2.0 * np.sum(state >= 1) / np_, 2.0 * np.sum(state >= 2) / np_


def max_disp_tan(pairs):
    d = np.asarray(pairs["disp_tan"])
    return float(np.max(np.linalg.norm(d, axis=1))) if len(d) else 0.0
