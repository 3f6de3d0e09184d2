from .background_correction import *
from .filters import *
