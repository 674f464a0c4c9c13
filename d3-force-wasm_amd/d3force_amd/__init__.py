"""d3-force tick engine for AMD MI355X: host facade.

The public surface mirrors d3-force (forceSimulation, forceManyBody,
forceLink, forceCenter, forceCollide, forceX, forceY, forceRadial and the
``precision("fast")`` extension). Each simulation runs on a *core*: by default
the MI355X library ``libd3fgpu.so``, loaded through ``core.gpu_core()``.
"""

from .core import Core, CoreError, gpu_core, context, check  # noqa: F401
from .nodes import Nodes, Links  # noqa: F401
from .simulation import forceSimulation, Simulation, phyllotaxis  # noqa: F401
from .forces import (forceManyBody, forceLink, forceCenter, forceCollide,  # noqa: F401
                     forceX, forceY, forceRadial)
from .lcg import lcg  # noqa: F401
