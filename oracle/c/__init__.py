"""C/OpenMP restatement of the alpacc runtime (test infrastructure and CPU baseline)."""
