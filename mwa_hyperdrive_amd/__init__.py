"""MI355X-native sky-model visibility predictor (drop-in for hyperdrive's SkyModeller)."""

from .modeller import AnalyticBeam, NoBeam, SkyModeller, new_sky_modeller  # noqa: F401
from .srclist import (  # noqa: F401
    CurvedPowerLaw,
    FluxDensity,
    FluxList,
    Gaussian,
    Point,
    PowerLaw,
    Shapelet,
    Source,
    SourceComponent,
    SourceList,
)
