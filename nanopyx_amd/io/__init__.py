from .tiff import imread, imwrite

__all__ = ["imread", "imwrite"]
