"""Particle property schema (reference: src/particles.py:1-162).

The names are the data format shared with the reference (and with PySPH
output files), so they are kept exactly."""

BASE_PROPERTIES = (
    "x", "y", "z", "gid", "tag",
    {"name": "itype", "type": "int", "default": 1},
    "u", "v", "w", "rho", "rho_sum", "h", "m", "cs", "p", "e",
    "au", "av", "aw", "arho", "ae", "ah",
    "x0", "y0", "z0", "u0", "v0", "w0",
    "upred", "vpred", "wpred", "epred", "e0", "h0",
    "div", "dt_cfl", "dt_force", "dt_adapt",
    "grhox", "grhoy", "grhoz", "dwdh", "omega", "converged",
    "alpha1", "alpha10", "Bx", "By", "Bz",
)

OUTPUT_PROPERTIES = (
    "x", "y", "z", "gid", "u", "v", "w", "rho", "h", "m", "itype", "alpha1", "p", "e", "Bx", "By", "Bz",
)

_STATE_PROPERTIES = ("Bevolx", "Bevoly", "Bevolz", "psi")

_INTERNAL_PROPERTIES = (
    "Bevolxpred", "Bevolypred", "Bevolzpred", "psipred", "Bxpred", "Bypred", "Bzpred",
    "aBevolx", "aBevoly", "aBevolz", "apsi", "Bevolx0", "Bevoly0", "Bevolz0", "psi0",
    "divBsymm", "divBdiff", "div_for_psi",
)

_CULLEN_DEHNEN_PROPERTIES = (
    "alphaloc", "xilimiter", "divvdt",
    "cd_rxx", "cd_rxy", "cd_rxz", "cd_ryy", "cd_ryz", "cd_rzz",
    "cd_vxx", "cd_vxy", "cd_vxz", "cd_vyx", "cd_vyy", "cd_vyz", "cd_vzx", "cd_vzy", "cd_vzz",
    "cd_axx", "cd_axy", "cd_axz", "cd_ayx", "cd_ayy", "cd_ayz", "cd_azx", "cd_azy", "cd_azz",
)

_CURRENT_STRAIN_PROPERTIES = (
    "dvdx_xx", "dvdx_xy", "dvdx_xz", "dvdx_yx", "dvdx_yy", "dvdx_yz", "dvdx_zx", "dvdx_zy", "dvdx_zz",
)

PROPERTIES = _STATE_PROPERTIES + _INTERNAL_PROPERTIES + _CULLEN_DEHNEN_PROPERTIES + _CURRENT_STRAIN_PROPERTIES

INITIAL_PREDICTOR_PAIRS = (
    ("Bevolxpred", "Bevolx"), ("Bevolypred", "Bevoly"), ("Bevolzpred", "Bevolz"), ("psipred", "psi"),
    ("upred", "u"), ("vpred", "v"), ("wpred", "w"), ("epred", "e"),
)

INITIAL_STATE_GPU_PROPERTIES = (
    "Bx", "By", "Bz", *_STATE_PROPERTIES, *(predicted for predicted, _current in INITIAL_PREDICTOR_PAIRS),
)
