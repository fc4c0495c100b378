"""Radar postprocessing restated on the CPU -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the library's radar postprocess node. It restates, in plain
Python and numpy:
  energy      src/gpu/nodeKernels.cu:82-101 (reflected polarisation) and :209-320 (kRadarComputeEnergy:
              physical-optics scattered field of one ray, BU, BR and the area factor), in FP64 complex arithmetic;
  clustering  src/graph/RadarPostprocessPointsNode.cpp:74-117 (greedy assignment in point order, then repeated
              merging of the first mergeable pair) and :302-398 (RadarCluster), with every comparison evaluated in
              FP32 like the reference's float code, including its merge rule for the radial-speed upper bound (the
              smaller of the two bounds is kept, :380-383);
  per cluster src/graph/RadarPostprocessPointsNode.cpp:119-174 (directional centre, RCS, received power).
The noise term of POWER / NOISE / SNR is the library's seeded normal stream, not the reference's unseeded one.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32
C0 = 299792458.0


def _reflect_pol(pol, n, d):
    """src/gpu/nodeKernels.cu:82-101."""
    if abs(float(np.dot(d, n))) == 1.0:
        return -pol
    pu = np.cross(d, n)
    pu = pu / np.linalg.norm(pu)
    pr = np.cross(d, pu)
    pr = pr / np.linalg.norm(pr)
    ref = d - n * (2.0 * np.dot(d, n))
    ref = ref / np.linalg.norm(ref)
    rpu = -pu
    rpr = np.cross(ref, rpu)
    return -np.dot(pol, pr) * rpr + np.dot(pol, pu) * rpu


def _compose(a, b):
    out = np.zeros((3, 4))
    out[:, :3] = a[:, :3] @ b[:, :3]
    out[:, 3] = a[:, :3] @ b[:, 3] + a[:, 3]
    return out


def energy(ray_pose, look_at, distance, normal, azimuth_step, elevation_step, frequency):
    """Per point (BU, BR, factor) as complex128 arrays (src/gpu/nodeKernels.cu:209-320)."""
    lam = C0 / float(frequency)
    k = 2.0 * math.pi / lam
    la = np.asarray(look_at, np.float64)
    n_pts = len(distance)
    bu = np.zeros(n_pts, np.complex128)
    br = np.zeros(n_pts, np.complex128)
    fac = np.zeros(n_pts, np.complex128)
    for p in range(n_pts):
        pose = _compose(la, np.asarray(ray_pose[p], np.float64))
        d = pose[:, 2] + pose[:, 3]   # pose * (0, 0, 1), a point transform
        pol = pose[:, 0] + pose[:, 3]  # pose * (1, 0, 0)
        nrm = la[:, :3] @ np.asarray(normal[p], np.float64)
        dist = float(distance[p])
        area = dist * dist * math.sin(elevation_step) * azimuth_step
        phi = math.atan2(d[1], d[2])
        the = math.acos(d[0] / np.linalg.norm(d))
        cp, sp, ct, st = math.cos(phi), math.sin(phi), math.cos(the), math.sin(the)
        dir_p = np.array([0.0, cp, -sp])
        dir_t = np.array([-st, sp * ct, cp * ct])
        refl = d - nrm * (2.0 * np.dot(d, nrm))
        refl = refl / np.linalg.norm(refl)
        rpol = _reflect_pol(pol, nrm, d)
        ap_e = np.exp(1j * k * dist) * rpol.astype(np.complex128)
        ap_h = -np.cross(ap_e, refl)
        f1 = -(np.cross(ap_e, -dir_p) + np.cross(ap_h, dir_t))
        f2 = -(np.cross(ap_e, dir_t) + np.cross(ap_h, dir_p))
        bu[p] = np.dot(f1, refl)
        br[p] = np.dot(f2, refl)
        fac[p] = 1j * (k * area * np.dot(refl, nrm) / (4.0 * math.pi)) * np.exp(-1j * k * dist)
    return bu, br, fac


def _scope_of(scopes, d):
    for s in scopes:
        if F(s[0]) <= d <= F(s[1]):
            return s
    return None


class _Cluster:
    def __init__(self, i, d, a, r, e):
        self.idx = [i]
        self.d = [d, d]
        self.a = [a, a]
        self.r = [r, r]
        self.e = [e, e]

    def add(self, i, d, a, r, e):
        self.idx.append(i)
        self.d = [min(self.d[0], d), max(self.d[1], d)]
        self.a = [min(self.a[0], a), max(self.a[1], a)]
        self.r = [r if np.isnan(self.r[0]) else min(self.r[0], r), r if np.isnan(self.r[1]) else max(self.r[1], r)]
        self.e = [min(self.e[0], e), max(self.e[1], e)]

    def candidate(self, d, a, r, s):
        ds, rs, az = F(s[2]), F(s[3]), F(s[4])
        ok = (d + ds >= self.d[0]) and (d - ds <= self.d[1]) and (a + az >= self.a[0]) and (a - az <= self.a[1])
        r_any_nan = np.isnan(self.r[0]) or np.isnan(self.r[1])
        return ok and (r_any_nan or ((r + rs >= self.r[0]) and (r - rs <= self.r[1])))

    def can_merge(self, o, scopes):
        def overlap(a, b):
            return (b[0] <= a[1] and b[1] >= a[0]) or (a[0] <= b[1] and a[1] >= b[0])

        def near(a, b, t):
            return abs(F(a[1] - b[0])) <= t or abs(F(b[1] - a[0])) <= t

        lo = max(self.d[0], o.d[0])
        hi = min(self.d[1], o.d[1])
        s = _scope_of(scopes, max(lo, hi))
        dist_ok = overlap(self.d, o.d) or near(self.d, o.d, F(s[2]))
        az_ok = overlap(self.a, o.a) or near(self.a, o.a, F(s[4]))
        r_ok = overlap(self.r, o.r) or near(self.r, o.r, F(s[3]))
        r_nan = np.isnan(self.r[0]) or np.isnan(self.r[1]) or np.isnan(o.r[0]) or np.isnan(o.r[1])
        return dist_ok and az_ok and (r_ok or r_nan)

    def take(self, o):
        self.d = [min(self.d[0], o.d[0]), max(self.d[1], o.d[1])]
        self.a = [min(self.a[0], o.a[0]), max(self.a[1], o.a[1])]
        self.r = [self.r[0] if np.isnan(o.r[0]) else min(o.r[0], self.r[0]),
                  self.r[1] if np.isnan(o.r[1]) else min(o.r[1], self.r[1])]  # the reference keeps the smaller
        self.e = [min(self.e[0], o.e[0]), max(self.e[1], o.e[1])]
        self.idx += o.idx


def _nan_min(a, b):
    # std::min(a, b) returns a unless b < a (NaN comparisons are false).
    return b if b < a else a


def cluster(distance, azimuth, radial_speed, elevation, scopes):
    """Clusters in creation order after merging; each has .idx (point indices in the reference's list order) and FP32
    bounds. `scopes` rows are (begin, end, distance_threshold, radial_speed_threshold, azimuth_threshold)."""
    d_, a_, r_, e_ = (np.asarray(x, np.float32) for x in (distance, azimuth, radial_speed, elevation))
    clusters: list[_Cluster] = []
    for i in range(len(d_)):
        d, a, r, e = d_[i], a_[i], r_[i], e_[i]
        s = _scope_of(scopes, d)
        if s is None:
            continue
        for c in clusters:
            if c.candidate(d, a, r, s):
                c.add(i, d, a, r, e)
                break
        else:
            clusters.append(_Cluster(i, d, a, r, e))
    done = False
    while len(clusters) > 1 and not done:
        done = True
        for i in range(len(clusters)):
            for j in range(i + 1, len(clusters)):
                if clusters[i].can_merge(clusters[j], scopes):
                    clusters[i].take(clusters[j])
                    del clusters[j]
                    done = False
                    break
            if not done:
                break
    return clusters


def centre_index(c, azimuth, elevation):
    """src/graph/RadarPostprocessPointsNode.cpp:379-398: the first point (list order) closest to the bounds' centre."""
    az, el = np.asarray(azimuth, np.float32), np.asarray(elevation, np.float32)
    ma = F((c.a[0] + c.a[1]) / F(2))
    me = F((c.e[0] + c.e[1]) / F(2))
    best, arg = np.inf, c.idx[0]
    for i in c.idx:
        v = F(abs(F(az[i] - ma)) + abs(F(el[i] - me)))
        if v < best:
            best, arg = v, i
    return arg


def postprocess(ray_pose, look_at, distance, azimuth, elevation, radial_speed, normal, scopes, azimuth_step,
                elevation_step, frequency, power_transmitted, device_gain):
    """Clusters, their centre indices, RCS (dBsm) and received power without noise (dBm)."""
    bu, br, fac = energy(ray_pose, look_at, distance, normal, azimuth_step, elevation_step, frequency)
    cl = cluster(distance, azimuth, radial_speed, elevation, scopes)
    lam = C0 / float(frequency)
    lam_db = 10.0 * math.log10(lam * lam)
    centres, rcs, power = [], [], []
    for c in cl:
        au = complex(np.sum(bu[c.idx] * fac[c.idx]))
        ar = complex(np.sum(br[c.idx] * fac[c.idx]))
        r = 10.0 * math.log10(4.0 * math.pi * (abs(au) ** 2 + abs(ar) ** 2))
        ci = centre_index(c, azimuth, elevation)
        d = float(distance[ci])
        mult = 10.0 * math.log10((4.0 * math.pi) ** 3) + 10.0 * math.log10(d ** 4)
        centres.append(ci)
        rcs.append(r)
        power.append(power_transmitted + 2.0 * device_gain + r + lam_db - mult)
    return cl, np.array(centres, np.int64), np.array(rcs), np.array(power)
