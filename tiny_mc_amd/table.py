"""The shell table of tiny_mc.c:58-66, computed from heat / heat2 in double.

Same columns and printf formats as the reference (radius in microns, heat in
W/cm^3 for a 1 W source, the reference's error estimate), and the same
'# extra' line for the overflow shell."""
import math


def table_rows(heat, heat2, photons, microns_per_shell):
    """(radius, heat, error) per shell 0..S-2, and the extra value."""
    S = len(heat)
    t = 4.0 * math.pi * microns_per_shell ** 3 * photons / 1e12
    rows = []
    for i in range(S - 1):
        vol = i * i + i + 1.0 / 3.0
        var = heat2[i] - heat[i] * heat[i] / photons
        err = math.sqrt(var) / t / vol if var >= 0 else float("nan")
        rows.append((i * microns_per_shell, heat[i] / t / vol, err))
    return rows, heat[S - 1] / photons


def format_table(heat, heat2, photons, microns_per_shell):
    rows, extra = table_rows(heat, heat2, photons, microns_per_shell)
    out = ["# Radius\tHeat", "# [microns]\t[W/cm^3]\tError"]
    out += ["%6.0f\t%12.5f\t%12.5f" % r for r in rows]
    out.append("# extra\t%12.5f" % extra)
    return "\n".join(out) + "\n"
