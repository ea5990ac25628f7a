"""MI355X-native BEM mat-vec (libBEEP hot path) -- see include/beep_hip.h."""
from .scene import Mesh, Scene
