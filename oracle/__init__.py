"""CPU oracle and CPU-baseline code for the synchrosqueezing path.  Test and
measurement infrastructure only; never imported by ssqueezepy_amd."""
