"""Node recorder rules restated — TEST INFRASTRUCTURE.

Only tests/ import this module; the product package never does. It restates the reference's
Node recorder (src/exec/recorders.cpp, after OpenSees' NodeRecorder.cpp and DataFileStream.cpp):

  - due (recorders.cpp:212-213): after a commit at `time`, a recorder with interval dT records
    when dT == 0 or time - next >= -dT * rtol (rtol 1e-5 for Node recorders), and then
    next = time + dT; next starts at 0;
  - line (recorders.cpp:315-325): the time (with -time), then every value, each printed as
    %.{precision}g, joined by one space, ending in a newline.

Pinned by the reference's own recorder output: the time column of
tests/reference/data/system/column25_ssp_explicit/opensees_disp.txt (dt 2.5e-4, -dT 0.01, 200
records) and the first lines of column4_pimy_std_implicit/opensees_disp.txt (precision 17, fixed
DOFs printed 0), kept in tests/golden/recorder_text.json.
"""


def due_indices(times, dT, rtol=1e-5):
    """Indices of the committed `times` at which a recorder of interval dT records."""
    out, nxt = [], 0.0
    for i, t in enumerate(times):
        if dT == 0.0 or t - nxt >= -dT * rtol:
            if dT != 0.0:
                nxt = t + dT
            out.append(i)
    return out


def line(time, values, precision=6, echo_time=True):
    """One record's text line (with its newline)."""
    toks = [f"{time:.{precision}g}"] if echo_time else []
    toks += [f"{v:.{precision}g}" for v in values]
    return " ".join(toks) + "\n"


def accumulated_times(dt, nsteps, t0=0.0):
    """The committed times of nsteps steps of dt from t0, accumulated as the integrators do
    (time += dt per step)."""
    out, t = [], t0
    for _ in range(nsteps):
        t += dt
        out.append(t)
    return out
