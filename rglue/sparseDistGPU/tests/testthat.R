library(testthat)
library(sparseDistGPU)
test_check("sparseDistGPU")
