## A request loop over a resident reference Y (sd_server_* in include/sd.h).
## Each request equals sparseKNN(X, Y, k, method, dist, drop_degenerate=)
## bit for bit; snn = TRUE adds the shared-neighbour graph of its columns.

gpuServer <- function(Y, slots = 4L, chunk_cols = 4096L, budget_bytes = 0, max_pending = 0) {
  .server_create(.as_dgc(Y), .chk_count(slots, "slots"), .chk_count(chunk_cols, "chunk_cols"),
                 budget_bytes, max_pending)
}

gpuSubmit <- function(server, X, k = 10, method = "binary", dist = TRUE,
                      drop_degenerate = TRUE, snn = FALSE, prune = 0) {
  method <- match.arg(method, .methods)
  .check_direction(method, dist)
  X <- .as_dgc(X)
  structure(.server_submit(server, X, .chk_count(k, "k"), .method_code(method), dist,
                           drop_degenerate, snn, prune),
            ncol = ncol(X), colnames = colnames(X))
}

gpuResult <- function(server, ticket) {
  res <- .server_result(server, as.numeric(ticket), attr(ticket, "ncol"))
  rownames(res$idx) <- rownames(res$dist) <- attr(ticket, "colnames")
  res
}

gpuServerClose <- function(server) invisible(.server_close(server))

gpuAutotune <- function(measure = TRUE) .autotune(measure)
