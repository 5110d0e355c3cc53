"""``python -m stencil_benchmarks_amd ...`` is the ``sbench`` command."""

from .cli import main

if __name__ == "__main__":
    main()
