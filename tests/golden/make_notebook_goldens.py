"""Extract the reference-held golden values for the Taylor hot path into a fixture.

The executed outputs of the reference's documentation notebooks are the only
absolute numbers the reference holds for the batch/ensemble integrator. This
script reads those notebooks as JSON text (no reference code is imported or run),
pulls the printed numbers out of specific cells and writes them, with their
source file:line, to ``notebook_goldens.json`` beside this script.

Run it in a checkout that has the reference at /root/reference:
    python tests/golden/make_notebook_goldens.py
The tests only read the JSON it wrote.
"""

import json
import os
import re

NB_DIR = "/root/reference/doc/notebooks"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_goldens.json")

_FLOAT = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _cells(nb):
    with open(os.path.join(NB_DIR, nb)) as f:
        return json.load(f)["cells"]


def _line_of(nb, needle):
    with open(os.path.join(NB_DIR, nb)) as f:
        for i, line in enumerate(f, 1):
            if needle in line:
                return i
    raise KeyError(f"{needle!r} not found in {nb}")


def _out_text(cell):
    parts = []
    for o in cell.get("outputs", []):
        if "text" in o:
            parts.append("".join(o["text"]))
        elif "data" in o and "text/plain" in o["data"]:
            parts.append("".join(o["data"]["text/plain"]))
    return "\n".join(parts)


def _cell_by_source(nb, marker, nth=0):
    hits = [c for c in _cells(nb) if c["cell_type"] == "code" and marker in "".join(c["source"])]
    return hits[nth]


def _floats(text):
    return [float(x) for x in _FLOAT.findall(text)]


def main():
    g = {}
    ai = "The adaptive integrator.ipynb"
    cu = "Customising the adaptive integrator.ipynb"
    bm = "Batch mode overview.ipynb"
    vo = "var_ode_sys.ipynb"

    # --- Pendulum x' = v, v' = -9.8 sin x, state [0.05, 0.025], tol = eps (order 20) ---
    g["pend_first_step"] = {
        "src": f"doc/notebooks/{ai}:118,159-160",
        "h": 0.21605277478009474,
        "state": [0.04399644836992638, -0.07844245547068798],
    }
    g["pend_step_backward"] = {"src": f"doc/notebooks/{ai}:192", "h": -0.21312300047513288}
    g["pend_clamped"] = {"src": f"doc/notebooks/{ai}:227-231", "h": [0.01, -0.02], "outcome": "time_limit"}
    txt = _out_text(_cell_by_source(ai, 'print("Current time        : {}".format(ta.time))'))
    fl = _floats(txt)
    g["pend_after_clamped"] = {
        "src": f"doc/notebooks/{ai}:{_line_of(ai, '-0.007070225695038143')}-{_line_of(ai, '0.04981102')}",
        "time": fl[0],
        "state": fl[1:3],
        "state_digits": 8,
    }
    txt = _out_text(_cell_by_source(ai, "status, min_h, max_h, nsteps, _, _ = ta.propagate_for(delta_t=5.0)"))
    fl = _floats(txt)
    g["pend_propagate_for_5"] = {"src": f"doc/notebooks/{ai}:374-377", "min_h": fl[0], "max_h": fl[1], "nsteps": int(fl[2]), "time": fl[3]}
    g["pend_propagate_until_20"] = {"src": f"doc/notebooks/{ai}:380-383", "min_h": fl[4], "max_h": fl[5], "nsteps": int(fl[6]), "time": fl[7]}
    txt = _out_text(_cell_by_source(ai, "# Propagate back to t = 0."))
    fl = _floats(txt)
    g["pend_propagate_until_0"] = {
        "src": f"doc/notebooks/{ai}:440-452",
        "min_h": fl[0], "max_h": fl[1], "nsteps": int(fl[2]), "time": fl[3],
        "state": [0.050000000000000044, 0.02499999999999999],
    }
    assert 0.050000000000000044 in _floats(txt) and 0.02499999999999999 in _floats(txt)
    txt = _out_text(_cell_by_source(ai, "ta.propagate_until(t=0.5, max_delta_t=0.1, callback=cb);"))
    g["pend_callback_times"] = {"src": f"doc/notebooks/{ai}:{_line_of(ai, 'Current time: 0.30000000000000004')}", "times": _floats(txt)}
    txt = _out_text(_cell_by_source(ai, "out = ta.propagate_grid(grid=[0.0, 0.1"))
    fl = _floats(txt)
    # (outcome, min_h, max_h, nsteps, None, array[11][2]) -- numbers after the outcome code.
    g["pend_grid"] = {
        "src": f"doc/notebooks/{ai}:606-669",
        "min_h": fl[1], "max_h": fl[2], "nsteps": int(fl[3]),
        "grid": [0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0],
        "out": [fl[4 + 2 * i: 6 + 2 * i] for i in range(11)],
        "out_digits": 8,
    }
    # tol = 1e-9: order 12, forth-and-back round trip.
    g["pend_tol1e9"] = {
        "src": f"doc/notebooks/{cu}:51,110",
        "order": 12,
        "state": [0.05000000000131285, 0.02499999999755865],
    }
    assert "Taylor order            : 12" in _out_text(_cell_by_source(cu, "tol=1e-9"))

    # --- Forced damped pendulum, batch of 4 ---------------------------------------------
    fdp = {"state": [[0.00, 0.01, 0.02, 0.03], [1.85, 1.86, 1.87, 1.88]], "pars": [[0.10, 0.11, 0.12, 0.13]]}
    g["fdp_setup"] = dict(fdp, src=f"doc/notebooks/{bm}:132,167-169", sys="x'=v, v'=cos(t)-p0*v-sin(x)")
    txt = _out_text(_cell_by_source(bm, "ta.step()\n"))
    fl = _floats(txt)
    g["fdp_step"] = {"src": f"doc/notebooks/{bm}:242-245", "outcome": int(fl[0]), "h": [fl[1], fl[3], fl[5], fl[7]]}
    txt = _out_text(_cell_by_source(bm, 'print("State:\\n{}".format(ta.state))'))
    fl = _floats(txt)
    g["fdp_step_state"] = {"src": f"doc/notebooks/{bm}:282-287", "state": [fl[0:4], fl[4:8]], "time": fl[8:12], "digits": 8}
    txt = _out_text(_cell_by_source(bm, "ta.step([0.010, 0.011, 0.012, 0.013])"))
    fl = _floats(txt)
    g["fdp_clamped"] = {
        "src": f"doc/notebooks/{bm}:320-326",
        "outcome": int(fl[0]), "h": [fl[1], fl[3], fl[5], fl[7]],
        "state": [fl[8:12], fl[12:16]], "time": fl[16:20], "digits": 8,
    }
    txt = _out_text(_cell_by_source(bm, "ta.propagate_for([10.0, 11.0, 12.0, 13.0])"))
    fl = _floats(txt)
    res1 = [(int(fl[4 * i]), fl[4 * i + 1], fl[4 * i + 2], int(fl[4 * i + 3])) for i in range(4)]
    st1 = [fl[16:20], fl[20:24]]
    t1 = fl[24:28]
    res2 = [(int(fl[28 + 4 * i]), fl[29 + 4 * i], fl[30 + 4 * i], int(fl[31 + 4 * i])) for i in range(4)]
    st2 = [fl[44:48], fl[48:52]]
    t2 = fl[52:56]
    g["fdp_propagate_for"] = {"src": f"doc/notebooks/{bm}:378-382", "delta_t": [10.0, 11.0, 12.0, 13.0], "res": res1, "state": st1, "time": t1, "digits": 8}
    g["fdp_propagate_until"] = {"src": f"doc/notebooks/{bm}:386-391", "t": [20.0, 21.0, 22.0, 23.0], "res": res2, "state": st2, "time": t2, "digits": 8}
    txt = _out_text([c for c in _cells(bm) if "".join(c["source"]).strip() == "ta.propagate_res"][0])
    fl = _floats(txt)
    g["fdp_grid_res"] = {
        "src": f"doc/notebooks/{bm}:482-500",
        "grid": "np.repeat(np.linspace(0, 100, 1000), 4).reshape(1000, -1)",
        "res": [(int(fl[4 * i]), fl[4 * i + 1], fl[4 * i + 2], int(fl[4 * i + 3])) for i in range(4)],
    }
    txt = _out_text(_cell_by_source(bm, "ta.step(write_tc=True)\n\n# Print the array"))
    fl = _floats(txt)
    assert len(fl) == 2 * 21 * 4, len(fl)
    g["fdp_tc"] = {"src": f"doc/notebooks/{bm}:550-{_line_of(bm, '-1.90354790e-05')}", "tc": [[fl[(i * 21 + k) * 4:(i * 21 + k) * 4 + 4] for k in range(21)] for i in range(2)], "digits": 8}
    txt = _out_text(_cell_by_source(bm, "ta.update_d_output([0.01, 0.02, 0.03, 0.04])"))
    fl = _floats(txt)
    g["fdp_d_output"] = {"src": f"doc/notebooks/{bm}:{_line_of(bm, '0.01854043')}", "t": [0.01, 0.02, 0.03, 0.04], "out": [fl[0:4], fl[4:8]], "digits": 8}

    # --- Variational forced damped pendulum (order 2), state [0.2, 0.3], par 0.4 -------
    txt = _out_text(_cell_by_source(vo, "ta_var.propagate_until(3.0)\n\n# Print"))
    fl = _floats(txt)
    g["var2_state_t3"] = {"src": f"doc/notebooks/{vo}:361-363", "state": fl[:12], "digits": 8, "ic": [0.2, 0.3], "pars": [0.4], "t": 3.0}
    txt = _out_text(_cell_by_source(vo, "Multiindex: {ta_var.get_mindex(idx)}"))
    g["var2_order2"] = {
        "src": f"doc/notebooks/{vo}:529-534",
        "mindex": [[0, 2, 0], [0, 1, 1], [0, 0, 2], [1, 2, 0], [1, 1, 1], [1, 0, 2]],
        "values": [float(m) for m in re.findall(r"derivative value: ([-+0-9.eE]+)", txt)],
    }
    txt = _out_text(_cell_by_source(vo, 'print(f"Non-variational state: {ta.state}")'))
    fl = _floats(txt)
    g["var2_displaced"] = {"src": f"doc/notebooks/{vo}:650-651", "dx": 1e-4, "dv": -2e-4, "nonvar_state": fl[0:2], "digits": 8}

    # Continuous output of the forced damped pendulum batch, propagate_until([10, 10.1, 10.2, 10.3]).
    txt = _out_text([c for c in _cells(bm) if "".join(c["source"]).strip() == "c_arr[0, 1, 2]"][0])
    g["fdp_c_output"] = {"src": f"doc/notebooks/{bm}:{_line_of(bm, '1.8917796783928316')}",
                         "t_final": [10.0, 10.1, 10.2, 10.3],
                         "times": [[0.01, 0.02, 0.03, 0.04], [1.01, 1.02, 1.03, 1.04], [2.01, 2.02, 2.03, 2.04]],
                         "c_arr_0_1_2": _floats(txt)[0]}

    # Event detection (pendulum x' = v, v' = -9.8 sin(x), x0 = -0.05, v0 = 0).
    ed = "Event detection.ipynb"
    txt = _out_text(_cell_by_source(ed, "print(zero_vel_times)", 0))
    g["ev_zero_vel_times"] = {"src": f"doc/notebooks/{ed}:{_line_of(ed, '[0.0, 1.003701787940065')}",
                              "times": _floats(txt), "event": "nt v = 0, direction any", "ic": [-0.05, 0.0],
                              "t_end": 5.0}
    txt = _out_text(_cell_by_source(ed, "print(zero_vel_times)", 1))
    g["ev_positive_times"] = {"src": f"doc/notebooks/{ed}:{_line_of(ed, '[0.0, 2.00740357588013, 4.01480715176026]')}",
                              "times": _floats(txt), "event": "nt v = 0, direction positive", "ic": [-0.05, 0.0],
                              "t_end": 5.0}
    txt = _out_text(_cell_by_source(ed, "ta.propagate_until(5.0)"))
    trig = re.findall(r"Event (\d) triggering at t=([-+0-9.eE]+)", txt)
    tail = txt[txt.index("(<taylor_outcome"):]
    fl = _floats(tail.split(">", 1)[1])
    g["ev_two_events"] = {"src": f"doc/notebooks/{ed}:{_line_of(ed, 'Event 0 triggering at t=0.0')}",
                          "events": [[int(i), float(t)] for i, t in trig], "ev0": "v", "ev1": "v * v - 1e-12",
                          "ic": [-0.05, 0.0], "t_end": 5.0, "outcome": "time_limit", "min_h": fl[0], "max_h": fl[1],
                          "nsteps": int(fl[2])}
    txt = _out_text(_cell_by_source(ed, "x_hist = ta.propagate_grid(t_grid)[5][:, 0]", 0))
    g["ev_x_at_zero_vel"] = {"src": f"doc/notebooks/{ed}:{_line_of(ed, 'Value of x when v is zero: -0.05')}",
                             "x": [float(m) for m in re.findall(r"Value of x when v is zero: ([-+0-9.eE]+)", txt)]}

    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print(f"wrote {OUT} ({len(g)} goldens)")


if __name__ == "__main__":
    main()
