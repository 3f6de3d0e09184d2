from .io_ import *
