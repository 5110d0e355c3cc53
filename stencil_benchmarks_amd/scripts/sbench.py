"""The ``sbench`` console script: every registered benchmark as a
sub-command (``sbench stencils mi355x horizontal-diffusion fused ...``)."""

from stencil_benchmarks_amd.cli import main

if __name__ == "__main__":
    main()
