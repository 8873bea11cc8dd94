from .mgpcg import MGPCG  # noqa: F401
