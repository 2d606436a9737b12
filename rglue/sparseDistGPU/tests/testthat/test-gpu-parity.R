## Parity with the CPU package on seeded inputs: identical indices and
## distances (the GPU path follows the reference's FP64 operation order).
skip_if_not_installed("sparseDist")

set.seed(1)
X <- abs(Matrix::rsparsematrix(300, 400, density = 0.05))
Y <- abs(Matrix::rsparsematrix(300, 250, density = 0.05))

test_that("sparseKNN matches the CPU package", {
  for (m in c("binary", "cosine", "euclidean", "manhattan")) {
    g <- sparseDistGPU::sparseKNN(X, k = 7, method = m)
    c <- sparseDist::sparseKNN(X, k = 7, method = m)
    expect_identical(g$idx, c$idx)
    expect_identical(g$dist, c$dist)
  }
})

test_that("sparseKNNCross matches the CPU package", {
  g <- sparseDistGPU::sparseKNNCross(X, Y, k = 5, method = "cosine", dist = FALSE)
  c <- sparseDist::sparseKNNCross(X, Y, k = 5, method = "cosine", dist = FALSE)
  expect_identical(g, c)
})

test_that("sparseSNN matches the CPU package", {
  nn <- sparseDistGPU::sparseKNN(X, k = 10, method = "cosine")
  expect_identical(sparseDistGPU::sparseSNN(nn, prune = 1 / 15),
                   sparseDist::sparseSNN(nn, prune = 1 / 15))
})

test_that("errors carry the reference's text", {
  expect_error(sparseDistGPU::sparseKNN(X, k = 0), "'k' must be a positive integer")
  expect_error(sparseDistGPU::sparseKNN(X, method = "euclidean", dist = FALSE),
               "only meaningful")
})

test_that("a served request equals the one-shot call", {
  s <- sparseDistGPU::gpuServer(Y)
  t <- sparseDistGPU::gpuSubmit(s, X[, 1:50], k = 5, method = "cosine")
  r <- sparseDistGPU::gpuResult(s, t)
  sparseDistGPU::gpuServerClose(s)
  one <- sparseDistGPU::sparseKNN(X[, 1:50], Y, k = 5, method = "cosine")
  expect_identical(unname(r$idx), unname(one$idx))
})
