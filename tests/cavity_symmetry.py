"""The symmetries of the differentially heated cavity, as a measure on
(rho, u, v, w, T) fields; shared by the oracle pins and the GPU tests."""
import numpy as np


def cavity_asymmetry(fields, dim, t_sum=1.0):
    """Largest departures of (rho, u, v, w, T) from the cavity's symmetries,
    relative to each field's scale, as a dict.

    2D (heated x walls, gravity along y): the half-turn (x, y) -> (H-x, H-y)
    with T -> Thot + Tcold - T, under which u and v change sign and rho
    stays.  3D (heated y walls, gravity along z, the x walls adiabatic): the
    mirror x -> H-x (u changes sign), and the half turn about the x axis,
    (y, z) -> (H-y, H-z) with T -> Thot + Tcold - T (v and w change sign).
    The half turns map the problem onto itself only when the buoyancy is odd
    in T about Thot + Tcold = 0: the reference's force is g beta T rho
    (natural-convection_2dim_mono.f90:478-486), with no reference
    temperature subtracted.  Arrays are [k][j][i] / [j][i], i fastest."""
    rho, u, v, w, T = fields
    sc_u = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max() if dim == 3 else 0.0, 1e-300)
    out = {}
    if dim == 2:
        rot = lambda a: a[::-1, ::-1]   # noqa: E731
        out["half_turn"] = max(np.abs(rho - rot(rho)).max(),
                               np.abs(u + rot(u)).max() / sc_u,
                               np.abs(v + rot(v)).max() / sc_u,
                               np.abs(T - (t_sum - rot(T))).max())
    else:
        mir = lambda a: a[:, :, ::-1]   # noqa: E731
        rot = lambda a: a[::-1, ::-1, :]   # noqa: E731
        out["mirror_x"] = max(np.abs(rho - mir(rho)).max(), np.abs(u + mir(u)).max() / sc_u,
                              np.abs(v - mir(v)).max() / sc_u, np.abs(w - mir(w)).max() / sc_u,
                              np.abs(T - mir(T)).max())
        out["half_turn_yz"] = max(np.abs(rho - rot(rho)).max(), np.abs(u - rot(u)).max() / sc_u,
                                  np.abs(v + rot(v)).max() / sc_u, np.abs(w + rot(w)).max() / sc_u,
                                  np.abs(T - (t_sum - rot(T))).max())
    return out
