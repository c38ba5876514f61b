"""Model specs read by the TOML front end from the fixture model files under
golden/toml (written in BEEP's input format for these tests)."""
import os

from beep_amd import toml_spec

_DIR = os.path.join(os.path.dirname(__file__), "golden", "toml")


def _load(name, data):
    spec, values, _ = toml_spec.load(os.path.join(_DIR, name), datadir=os.path.join(_DIR, data))
    return spec, values


def toml_seir_erlang():
    """SEIR, Erlang(2) latent period, 2 steps per day, weekly counts and populations."""
    return _load("seir_erlang.toml", "data_seir_erlang")


def toml_seird_branch():
    """SEIRD, Erlang(3) infectious period, Dirichlet branches, log-smoothed 3-knot R(t)."""
    return _load("seird_branch.toml", "data_seird_branch")


def toml_sir_3areas():
    """SIR over three areas: geographic mixing, mixing spline, per-area and per-region counts."""
    return _load("sir_3areas.toml", "data_sir_3areas")


def toml_sird_ages():
    """SIRD with three age groups: contact matrix, age-dependent periods and branches, efoi by age."""
    return _load("sird_ages.toml", "data_sird_ages")


def toml_seir_sex_ages():
    """SEIR over two areas with two age groups and two sexes: susceptibility by age and sex, per-sex counts."""
    return _load("seir_sex_ages.toml", "data_seir_sex_ages")


def toml_sird_strains():
    """SIRD with two strains and two age groups: strain susceptibility and Rfactor, pooled susceptibles."""
    return _load("sird_strains.toml", "data_sird_strains")
