module gofft-hip/go/gofft

go 1.21
