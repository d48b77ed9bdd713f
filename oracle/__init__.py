"""Oracle package: CPU restatements used only as the checker (tests, bench cpu_baseline)."""
