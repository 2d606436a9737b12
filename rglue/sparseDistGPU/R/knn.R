## GPU versions of sparseKNN(), sparseKNNCross() and sparseSNN(): the
## reference's arguments, defaults, checks and result layout; the block loop
## is one call into libsd_hip.so.

sparseKNN <- function(X, Y = NULL, k = 10, method = "binary", dist = TRUE,
                      include_self = FALSE, drop_degenerate = TRUE,
                      block_size = 256L, ncores = 1, verbose = TRUE) {
  if (is.null(X)) stop("'X' must be provided.", call. = FALSE)
  .chk_flag(dist, "dist")
  .chk_flag(include_self, "include_self")
  .chk_flag(drop_degenerate, "drop_degenerate")
  .chk_flag(verbose, "verbose")
  .chk_count(ncores, "ncores", allow_zero = TRUE)
  k <- .chk_count(k, "k")
  block_size <- .chk_count(block_size, "block_size")
  method <- match.arg(method, .methods)
  .check_direction(method, dist)
  X <- .as_dgc(X)
  if (ncol(X) < 1L) stop("'X' must have at least one column.", call. = FALSE)
  self <- is.null(Y)
  if (!self) {
    Y <- .as_dgc(Y)
    if (nrow(X) != nrow(Y)) stop("'X' and 'Y' must have the same number of rows.", call. = FALSE)
  }
  res <- .knn_gpu(X, Y, k, .method_code(method), dist, include_self, drop_degenerate, block_size)
  if (res$warn)
    warning(sprintf("k (%d) exceeds the %d available neighbours; returning %d.", k, res$k,
                    res$k), call. = FALSE)
  idx <- res$idx
  dst <- res$dist
  rownames(idx) <- rownames(dst) <- colnames(X)
  out <- list(idx = idx, dist = dst)
  attr(out, "self_search") <- self
  out
}

sparseKNNCross <- function(X, Y, k = 10L,
                           method = c("binary", "jaccard", "cosine", "euclidean",
                                      "manhattan", "pearson", "js", "covariance"),
                           dist = TRUE, drop_degenerate = TRUE,
                           block_size = 256L, ncores = 1L, verbose = TRUE) {
  if (missing(Y) || is.null(Y))
    stop(paste0("'Y' is required. For a single matrix use sparseKNN(), which already ",
                "exploits symmetry via the full/diag arguments."), call. = FALSE)
  .chk_flag(dist, "dist")
  .chk_flag(drop_degenerate, "drop_degenerate")
  .chk_flag(verbose, "verbose")
  .chk_count(ncores, "ncores", allow_zero = TRUE)
  k <- .chk_count(k, "k")
  block_size <- .chk_count(block_size, "block_size")
  method <- match.arg(method)
  .check_direction(method, dist)
  X <- .as_dgc(X)
  Y <- .as_dgc(Y)
  if (nrow(X) != nrow(Y)) stop("'X' and 'Y' must have the same number of rows.", call. = FALSE)
  if (ncol(X) < 1L) stop("'X' must have at least one column.", call. = FALSE)
  if (ncol(Y) < 1L) stop("'Y' must have at least one column.", call. = FALSE)
  res <- .knn_cross_gpu(X, Y, k, .method_code(method), dist, drop_degenerate, block_size)
  if (bitwAnd(res$warn, 1L))
    warning(sprintf(paste0("k (%d) exceeds the %d available neighbours for the forward ",
                           "direction; returning %d."), k, res$k_forward, res$k_forward),
            call. = FALSE)
  if (bitwAnd(res$warn, 2L))
    warning(sprintf(paste0("k (%d) exceeds the %d available neighbours for the reverse ",
                           "direction; returning %d."), k, res$k_reverse, res$k_reverse),
            call. = FALSE)
  rownames(res$forward$idx) <- rownames(res$forward$dist) <- colnames(X)
  rownames(res$reverse$idx) <- rownames(res$reverse$dist) <- colnames(Y)
  list(forward = res$forward, reverse = res$reverse)
}

sparseSNN <- function(idx, include_self = TRUE, prune = 0,
                      symmetrise = TRUE, edges = c("knn", "shared"),
                      ncores = 1, verbose = TRUE) {
  sym_given <- !missing(symmetrise)
  if (is.list(idx)) {
    ss <- attr(idx, "self_search")
    if (isFALSE(ss))
      stop(paste0("sparseSNN() requires a self-search sparseKNN() result; the indices from ",
                  "sparseKNN(X, Y) refer to the columns of Y, which are a different node set."),
           call. = FALSE)
    if (!isTRUE(ss))
      stop(paste0("'idx' must be a matrix, or a list produced by sparseKNN(). This list ",
                  "carries no 'self_search' marker -- it may have been dropped by subsetting ",
                  "-- so it cannot be confirmed to describe a single node set. Pass the index ",
                  "matrix itself if that is intended."), call. = FALSE)
    idx <- idx$idx
  }
  if (!is.matrix(idx)) stop("'idx' must be a matrix or a sparseKNN() result.", call. = FALSE)
  if (nrow(idx) < 1L || ncol(idx) < 1L)
    stop("'idx' must have at least one row and one column.", call. = FALSE)
  if (!is.numeric(idx)) stop("'idx' must be a numeric or integer matrix.", call. = FALSE)
  ok <- is.na(idx) | (is.finite(idx) & idx == floor(idx) & idx >= 1 & idx <= nrow(idx))
  if (!all(ok))
    stop(paste0("'idx' entries must be NA or finite, whole-number, 1-based row indices in ",
                "[1, nrow(idx)]."), call. = FALSE)
  .chk_flag(include_self, "include_self")
  .chk_flag(symmetrise, "symmetrise")
  .chk_flag(verbose, "verbose")
  edges <- match.arg(edges)
  if (edges == "shared" && sym_given && !symmetrise)
    warning(paste0("'symmetrise' is ignored when edges = \"shared\": that edge set is ",
                   "symmetric by construction."), call. = FALSE)
  if (!is.numeric(prune) || length(prune) != 1L || is.na(prune) || !is.finite(prune) ||
      prune < 0 || prune > 1)
    stop("'prune' must be a finite number between 0 and 1.", call. = FALSE)
  .chk_count(ncores, "ncores", allow_zero = TRUE)
  storage.mode(idx) <- "integer"
  g <- .snn_gpu(idx, include_self, prune, symmetrise, if (edges == "knn") 0L else 1L)
  if (!is.null(rownames(idx))) dimnames(g) <- list(rownames(idx), rownames(idx))
  g
}
